/*
 * Temporally blocked FDTD + density tile kernel for gfx950.
 *
 * Geometry.  All runs of a context are laid out back to back in one
 * "super-grid" of G cells: run r occupies cells [base_r, base_r + N_r)
 * followed by one pseudo cell that holds the boundary value h[N_r] (always
 * zero).  A cell flagged `hzero` has its H forced to zero after every H
 * update (a run's first cell: h[0] = 0; the pseudo cell: h[N] = 0,
 * solver_cpu_fdtd.tpp:273-275), which decouples neighbouring runs exactly.
 *
 * One wavefront (64 lanes) owns one tile.  Lane l holds the C consecutive
 * cells g = g0 + l*C + c in registers, where g0 = t*owned - halo.  The
 * wavefront advances `nsteps <= halo` time steps without leaving the
 * registers; after s steps the outermost s cells at each tile edge are
 * stale, so only the `owned = 64*C - 2*halo` central cells are written
 * back.  Every cell's arithmetic is independent of the tiling, so results
 * do not depend on C or the temporal block.
 *
 * Step order inside the block is the plain one of solver_cpu_fdtd::run
 * (solver_cpu_fdtd.tpp:230-325): E update, sources, H update with
 * boundary condition, density update, record copy; with RED that of
 * solver_cpu_fdtd_red::run (solver_cpu_fdtd_red.tpp:451-549): density
 * update, E update, sources, H update with boundary condition, record copy.
 * Both orders have the same one-cell stencil reach per step, so the halo
 * rule above holds for either.
 */
#ifndef MBH_TILE_HPP
#define MBH_TILE_HPP

#include <type_traits>

#include "mbh_device.hpp"

namespace mbh {

constexpr int FDTD_STRIDE = 5; /* fac_a fac_b fac_c gamma dx_inv */
constexpr int BLOCH2_STRIDE = 28; /* two-level Cayley, Bloch form */

struct dev_source {
    uint64_t gx;         /* super-grid cell */
    const double* data;  /* n_t values */
    uint32_t type;
    uint32_t pad;
};

struct dev_record {
    uint64_t gbegin; /* super-grid cell of position_idx */
    uint64_t cols;
    uint64_t rows;
    double ratio;
    double* dst; /* scratch + offset_real */
    uint32_t type;
    uint32_t row_idx;
    uint32_t col_idx;
    uint32_t pad;
};

struct KArgs {
    const void* e_in;
    const void* h_in;
    const void* p_in;
    const void* d_in;
    void* e_out;
    void* h_out;
    void* p_out;
    void* d_out;
    uint64_t G;       /* cells in the super-grid */
    uint64_t Gown;    /* cells owned by tiles: G minus the trailing pseudo
                         cell (its H is read at step 0 only) */
    uint64_t dstride; /* elements between density components */
    const int32_t* mat;   /* [G], -1: pseudo cell */
    const uint8_t* hzero; /* [G] */
    const double* fdtd;   /* [nmat][FDTD_STRIDE] */
    const double* qm;     /* [nmat][qstride] */
    uint64_t qstride;
    const dev_source* src;     /* sorted by gx */
    const uint32_t* tile_src;  /* [2*num_tiles]: [lo, hi) */
    const dev_record* rec;
    const uint32_t* tile_rec_ptr; /* [num_tiles + 1] */
    const uint32_t* tile_rec_idx;
    const uint64_t* n_base; /* device: first step = *n_base + n_off */
    uint64_t n_off;
    uint32_t nsteps;
    uint32_t halo;
    uint32_t owned;
    uint32_t yshift; /* FP32 z-form storage holds y - y* (VALU and MFMA) */
    const void* frag;     /* MFMA form: [nmat][frag_stride bytes] */
    uint64_t frag_stride;
    const uint32_t* tile_list; /* tiles of this launch (NULL: blockIdx) */
    /* MFMA form: the first n_mixed tiles of the list hold more than one
     * material; tile_mat[i] (i: position in the list) is the material of
     * a uniform tile, plus TILE_MAT_QM when it has a density (NULL: the
     * kernel reads the cells' materials) */
    const int32_t* tile_mat;
    uint32_t n_mixed;
};
constexpr int32_t TILE_MAT_QM = 0x40000000;
/* kernel-table forms of the specialised kernels (internal, beside the
 * MBH_FORM_*): algo_noop_hz0, and algo_bloch2r with FP32 / FP64 state */
constexpr uint32_t KFORM_NOOP_HZ0 = 0x100;
constexpr uint32_t KFORM_BLOCH2R_MIXED = 0x101;
constexpr uint32_t KFORM_BLOCH2R_ZFORM64 = 0x102;

template<typename ST>
__device__ __forceinline__ double ld(const void* p, uint64_t i)
{
    return (double)(static_cast<const ST*>(p)[i]);
}
template<typename ST>
__device__ __forceinline__ void st(void* p, uint64_t i, double v)
{
    static_cast<ST*>(p)[i] = (ST)v;
}

/* read-only tables through the constant address space (scalar loads for
 * uniform indices) */
template<typename T>
__device__ __forceinline__ const __attribute__((address_space(4))) T*
as_const(const T* p)
{
    return (const __attribute__((address_space(4))) T*)p;
}

/* Whole-wavefront neighbour exchange with DPP: wave_shl:1 gives lane i the
 * value of lane i+1, wave_shr:1 the value of lane i-1 (the edge lane gets
 * 0; edge cells are halo cells).  Register-to-register, no LDS round trip. */
__device__ __forceinline__ double from_right(double x)
{
    const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(x), 0x130,
                                               0xf, 0xf, false);
    const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(x), 0x130,
                                               0xf, 0xf, false);
    return __hiloint2double(hi, lo);
}
__device__ __forceinline__ double from_left(double x)
{
    const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(x), 0x138,
                                               0xf, 0xf, false);
    const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(x), 0x138,
                                               0xf, 0xf, false);
    return __hiloint2double(hi, lo);
}

/* Runs `fn(material)` for every lane whose `m >= 0`, one distinct material
 * at a time; inside fn the material index is wave-uniform (SGPR), so
 * constant reads through it become scalar loads. */
template<typename F>
__device__ __forceinline__ void for_each_material(int32_t m, F&& fn)
{
    bool todo = m >= 0;
    for (;;) {
        unsigned long long mask = __ballot(todo);
        if (mask == 0ull) {
            break;
        }
        int leader = __builtin_ctzll(mask);
        int32_t mm = __builtin_amdgcn_readlane(m, leader);
        if (todo && m == mm) {
            fn(mm);
            todo = false;
        }
    }
}

__device__ __forceinline__ dev_record load_record(const dev_record* base,
                                                 uint32_t i)
{
    const auto* r = as_const(base) + i;
    dev_record R;
    R.gbegin = r->gbegin;
    R.cols = r->cols;
    R.rows = r->rows;
    R.ratio = r->ratio;
    R.dst = r->dst;
    R.type = r->type;
    R.row_idx = r->row_idx;
    R.col_idx = r->col_idx;
    R.pad = 0;
    return R;
}

/* copy_list_entry_dev::hasto_record (copy_list_entry.hpp:74-80) plus the
 * row bound */
__device__ __forceinline__ bool record_now(double ratio, uint64_t rows,
                                           uint64_t n)
{
    bool now = (n == 0);
    if (!now) {
        now = floor(ratio * (double)(n - 1)) != floor(ratio * (double)n);
    }
    return now && (uint64_t)floor(ratio * (double)n) < rows;
}

template<class A>
__device__ __forceinline__ double record_value(const KArgs& k, uint32_t type,
                                               uint32_t row, int32_t m,
                                               double e, double h, double p,
                                               const typename A::state& d)
{
    cdp q = (cdp)k.qm + (uint64_t)(m < 0 ? 0 : m) * k.qstride;
    switch (type) {
    case MBH_REC_ELECTRIC:
        return e;
    case MBH_REC_POLAR_DT:
        return p;
    case MBH_REC_MAGNETIC:
        return h;
    case MBH_REC_INVERSION:
        return (m < 0 || k.qstride == 0) ? 0.0 : A::inversion(q, d);
    default: /* density: only populations are recorded */
        return (m < 0 || k.qstride == 0) ? 0.0 : A::population(q, d, row);
    }
}

/* FP32 storage of the z-form holds y - y* (y* at OFF_YEQ of the material's
 * blob): v += sign * y* for a cell of material m with a density */
template<class A>
__device__ __forceinline__ void add_shift(const KArgs& k, int32_t m,
                                          double* v, double sign)
{
    /* loads from material 0's blob where m < 0 (every blob is there), so
     * that none waits in a branch of its own */
    const double* q =
        (const double*)k.qm + (uint64_t)(m < 0 ? 0 : m) * k.qstride;
    double y[A::DC];
#pragma unroll
    for (int j = 0; j < A::DC; j++) {
        y[j] = q[A::OFF_YEQ + j];
    }
    const bool on = (m >= 0) && (q[0] != 0.0);
#pragma unroll
    for (int j = 0; j < A::DC; j++) {
        v[j] = on ? fma(sign, y[j], v[j]) : v[j];
    }
}

/* sources staged per launch in LDS (values for the block's steps) */
constexpr int TILE_SOURCES = 8;
constexpr int MAX_BLOCK = 64; /* largest temporal block */

/* One tile t: all steps of the temporal block for the tile's cells.
 * With W > 1 the tile spans W wavefronts of one workgroup, each holding a
 * contiguous run of 64 C cells; at the inner edges between them the
 * neighbour's boundary H (before the E update) and E (before the H
 * update) come through LDS every step, so only the tile's two outer edges
 * carry a halo (64 C W - 2 tb owned cells instead of 64 C - 2 tb). */
template<class A, int C, typename ST, bool RED = false, int W = 1>
__device__ __forceinline__ void tile_body(const KArgs& k, uint64_t t)
{
    static_assert(W == 1 || !RED, "multi-wavefront tiles: plain order only");
    static_assert(W == 1 || A::STAGE == 0,
                  "multi-wavefront tiles: no per-wavefront LDS staging");
    __shared__ double s_src[TILE_SOURCES][MAX_BLOCK];
    /* boundary values of each wavefront's run for its neighbours */
    __shared__ double x_h[W > 1 ? W : 1]; /* H of the first cell */
    __shared__ double x_e[W > 1 ? W : 1]; /* E of the last cell */

    const int lane = threadIdx.x & 63;
    const int wv = (W > 1) ? (int)(threadIdx.x >> 6) : 0;
    const int64_t own_lo = (int64_t)(t * k.owned);
    const int64_t own_hi =
        ((own_lo + (int64_t)k.owned) < (int64_t)k.Gown) ? own_lo + k.owned
                                                         : (int64_t)k.Gown;
    const int64_t gl =
        own_lo - (int64_t)k.halo + (int64_t)(wv * 64 + lane) * C;
    const uint64_t n0 = *as_const(k.n_base) + k.n_off;
    const uint32_t nsteps = k.nsteps;

    double e[C], h[C], p[C];
    double fa[C], fb[C], fc[C], fg[C], fx[C];
    typename A::state d[C];
    int32_t m[C];
    bool hz[C];

    /* The state loads of all cells first, each from a clamped address
     * (cell 0 where the cell lies outside the grid) and selected after,
     * then the loads that depend on the cells' materials: no load waits
     * in a branch of its own, and each group is in flight together. */
#pragma unroll
    for (int c = 0; c < C; c++) {
        const int64_t g = gl + c;
        const bool in = (g >= 0) && (g < (int64_t)k.G);
        const uint64_t gc = in ? (uint64_t)g : 0u;
        const int32_t mv = k.mat[gc];
        const uint8_t hv = k.hzero[gc];
        const double ev = ld<ST>(k.e_in, gc);
        const double hv2 = ld<ST>(k.h_in, gc);
        const double pv = (A::DC > 0) ? ld<ST>(k.p_in, gc) : 0.0;
        double dv[A::DC > 0 ? A::DC : 1];
#pragma unroll
        for (int j = 0; j < A::DC; j++) {
            dv[j] = ld<ST>(k.d_in, (uint64_t)j * k.dstride + gc);
        }
        m[c] = in ? mv : -1;
        hz[c] = in ? (hv != 0) : true;
        e[c] = in ? ev : 0.0;
        /* the trailing pseudo cell is never owned, so never written back:
         * its H is h[N] of the initial state at step 0 and zero after (the
         * boundary condition) */
        h[c] = (in && (g + 1 < (int64_t)k.G || n0 == 0)) ? hv2 : 0.0;
        /* no density: p keeps its initial zero (lindblad_noop) and is
         * neither read nor written */
        p[c] = (A::DC > 0 && in) ? pv : 0.0;
#pragma unroll
        for (int j = 0; j < A::DC; j++) {
            d[c].v[j] = in ? dv[j] : 0.0;
        }
    }
#pragma unroll
    for (int c = 0; c < C; c++) {
        if constexpr (A::SHIFT && std::is_same<ST, float>::value) {
            if (k.yshift) {
                add_shift<A>(k, m[c], d[c].v, 1.0);
            }
        }
        /* material 0's row where the cell has none, then zeroed */
        const double* f =
            k.fdtd + (uint64_t)(m[c] < 0 ? 0 : m[c]) * FDTD_STRIDE;
        const double f0 = f[0], f1 = f[1], f2 = f[2], f3 = f[3], f4 = f[4];
        const bool known = m[c] >= 0;
        fa[c] = known ? f0 : 0.0;
        fb[c] = known ? f1 : 0.0;
        fc[c] = known ? f2 : 0.0;
        fg[c] = known ? f3 : 0.0;
        fx[c] = known ? f4 : 0.0;
        if constexpr (A::FAST_FDTD) {
            /* folded constants: e' = fa e + fg' p + fx' (h_r - h) with
             * fg' = -fb fg, fx' = fb fx (zero outside the grid, so those
             * cells stay zero); h' = hm h + fc' (e_c - e_l) with hm = 0,
             * fc' = 0 where H is held at zero */
            fg[c] = -(fb[c] * fg[c]);
            fx[c] = fb[c] * fx[c];
            fb[c] = hz[c] ? 0.0 : 1.0;
            fc[c] = hz[c] ? 0.0 : fc[c];
        }
    }

    /* material with a density, per cell (-1: vacuum or outside the grid),
     * and m1, the first of them in the tile. The lanes whose cells all have
     * m1 update with m1's constants as in a tile of one material (a tile
     * at an interface between m1 and vacuum is then as fast as one inside
     * the medium); the other lanes of the tile (a cell of another density
     * material, or one lane straddling the interface) take one update per
     * material. The arithmetic of a cell is the same either way. */
    int32_t ma[C];
    int32_t first = -1;
#pragma unroll
    for (int c = 0; c < C; c++) {
        bool act = false;
        if constexpr (A::DC > 0) {
            /* material 0's flag read where the cell has none */
            const bool q = A::has_qm(
                (cdp)k.qm + (uint64_t)(m[c] < 0 ? 0 : m[c]) * k.qstride);
            act = (m[c] >= 0) && q;
        }
        ma[c] = act ? m[c] : -1;
    }
#pragma unroll
    for (int c = C - 1; c >= 0; c--) {
        first = (ma[c] >= 0) ? ma[c] : first;
    }
    const unsigned long long act_lanes = __ballot(first >= 0);
    const int32_t m1 = act_lanes
        ? __builtin_amdgcn_readlane(first, __builtin_ctzll(act_lanes))
        : -1;
    bool all1 = true;
#pragma unroll
    for (int c = 0; c < C; c++) {
        all1 = all1 && (ma[c] == m1);
    }
    all1 = all1 && (m1 >= 0);
    const bool full1 = __all(all1);
    cdp q0 = (cdp)k.qm + (uint64_t)(m1 < 0 ? 0 : m1) * k.qstride;
    const bool q0_active = m1 >= 0;
    /* doubles of the material blob the algorithm wants staged in LDS */
    constexpr int STAGE = A::STAGE;
    __shared__ double s_q[STAGE > 0 ? STAGE : 1];
    if constexpr (STAGE > 0) {
        if (q0_active) {
            for (int i = lane; i < STAGE; i += 64) {
                s_q[i] = q0[i];
            }
        }
        __syncthreads();
    }

    /* sources of this tile: their values for the block's steps go to LDS */
    const uint32_t s_lo = as_const(k.tile_src)[2 * t];
    const uint32_t s_hi = as_const(k.tile_src)[2 * t + 1];
    const uint32_t nsrc = s_hi - s_lo;
    const bool src_staged = nsrc <= TILE_SOURCES && nsteps <= MAX_BLOCK;
    if (src_staged && nsrc > 0) {
        if (wv == 0) {
            for (uint32_t j = 0; j < nsrc; j++) {
                const double* data = as_const(k.src)[s_lo + j].data;
                s_src[j][lane] = (lane < (int)nsteps) ? data[n0 + lane] : 0.0;
            }
        }
        __syncthreads();
    }
    if constexpr (W > 1) {
        if (lane == 0) {
            x_h[wv] = h[0];
        }
        __syncthreads();
    }

    /* steps of this block at which any of the tile's records is due */
    const uint32_t r_lo = as_const(k.tile_rec_ptr)[t];
    const uint32_t r_hi = as_const(k.tile_rec_ptr)[t + 1];
    uint64_t rec_steps = 0;
    for (uint32_t ri = r_lo; ri < r_hi; ri++) {
        const uint32_t idx = as_const(k.tile_rec_idx)[ri];
        const double ratio = as_const(k.rec)[idx].ratio;
        const uint64_t rows = as_const(k.rec)[idx].rows;
        rec_steps |= __ballot(lane < (int)nsteps &&
                              record_now(ratio, rows, n0 + lane));
    }

    /* density update of every cell from its current field
     * (solver_cpu_fdtd.tpp:278-285, solver_cpu_fdtd_red.tpp:451-457) */
    auto density_update = [&]() {
        if constexpr (A::DC > 0) {
            auto update_m1 = [&]() {
                if constexpr (STAGE > 0) {
                    /* an offset the compiler cannot see through keeps the
                     * LDS reads inside the step (hoisted, the staged
                     * constants would occupy VGPRs for the whole loop) */
                    int off = 0;
                    asm volatile("" : "+s"(off));
                    A::template update_multi_p<C>((const double*)s_q + off,
                                                  d, e, p);
                } else {
                    A::template update_multi<C>(q0, d, e, p);
                }
            };
            /* small updates (two levels) get a second copy behind a
             * wave-uniform branch, so that tiles of one material run it
             * with full exec and no per-lane branch; for the large ones
             * the second copy costs more (instruction cache) than the
             * per-lane branch */
            constexpr bool TWO_COPIES = A::DC <= 4;
            if (TWO_COPIES && full1) {
                update_m1();
            } else if (all1) {
                update_m1();
            } else if (m1 >= 0) {
                /* lanes with a cell of another material: one update per
                 * material */
#pragma unroll
                for (int c = 0; c < C; c++) {
                    for_each_material(ma[c], [&](int32_t mm) {
                        A::update((cdp)k.qm + (uint64_t)mm * k.qstride, d[c],
                                  e[c], p[c]);
                    });
                }
            }
        }
    };

    for (uint32_t s = 0; s < nsteps; s++) {
        const uint64_t n = n0 + s;

        if constexpr (RED) {
            density_update();
        }

        /* E update (solver_cpu_fdtd.tpp:238-243) */
        double h_right = from_right(h[0]);
        if constexpr (W > 1) {
            /* the last lane's right neighbour is the next run's first
             * cell (beyond the tile's last run: halo, as with W = 1) */
            const double xr = (wv + 1 < W) ? x_h[wv + 1] : 0.0;
            h_right = (lane == 63) ? xr : h_right;
        }
        double en[C];
#pragma unroll
        for (int c = 0; c < C; c++) {
            const double hn = (c + 1 < C) ? h[(c + 1) % C] : h_right;
            if constexpr (A::FAST_FDTD && A::DC == 0) {
                en[c] = fma(fx[c], hn - h[c], fa[c] * e[c]); /* p = 0 */
            } else if constexpr (std::is_same<A, algo_bloch2r>::value) {
                /* lossless (fa = 1 in the grid; cells outside it and the
                 * pseudo cells hold E = 0 with fg = fx = 0) */
                en[c] = fma(fx[c], hn - h[c], fma(fg[c], p[c], e[c]));
            } else if constexpr (A::FAST_FDTD) {
                en[c] = fma(fx[c], hn - h[c], fma(fg[c], p[c], fa[c] * e[c]));
            } else {
                en[c] = (m[c] >= 0)
                    ? fa[c] * e[c] +
                        fb[c] * (-fg[c] * p[c] + (hn - h[c]) * fx[c])
                    : 0.0;
            }
        }

        /* sources, in list order (solver_cpu_fdtd.tpp:247-255) */
        if (nsrc > 0) {
            for (uint32_t j = 0; j < nsrc; j++) {
                const uint64_t gx = as_const(k.src)[s_lo + j].gx;
                const uint32_t type = as_const(k.src)[s_lo + j].type;
                const double v = src_staged
                    ? s_src[j][s]
                    : as_const(k.src)[s_lo + j].data[n];
                const int64_t off = (int64_t)gx - gl;
#pragma unroll
                for (int c = 0; c < C; c++) {
                    if (off == c) {
                        en[c] = (type == MBH_SRC_HARD) ? v : en[c] + v;
                    }
                }
            }
        }

        /* H update and boundary condition (solver_cpu_fdtd.tpp:263-275) */
        double e_left = from_left(en[C - 1]);
        if constexpr (W > 1) {
            if (lane == 63) {
                x_e[wv] = en[C - 1];
            }
            __syncthreads();
            const double xl = (wv > 0) ? x_e[wv - 1] : 0.0;
            e_left = (lane == 0) ? xl : e_left;
        }
#pragma unroll
        for (int c = 0; c < C; c++) {
            const double ep = (c > 0) ? en[(c + C - 1) % C] : e_left;
            if constexpr (std::is_same<A, algo_noop_hz0>::value ||
                          std::is_same<A, algo_bloch2r>::value) {
                /* every held cell starts at zero and its fc is zero */
                h[c] = fma(fc[c], en[c] - ep, h[c]);
            } else if constexpr (A::FAST_FDTD) {
                h[c] = fma(fc[c], en[c] - ep, fb[c] * h[c]);
            } else {
                h[c] = hz[c] ? 0.0 : h[c] + fc[c] * (en[c] - ep);
            }
        }
        if constexpr (W > 1) {
            /* the next step's E update reads it (every read of the old
             * value came before the barrier above) */
            if (lane == 0) {
                x_h[wv] = h[0];
            }
            __syncthreads();
        }

#pragma unroll
        for (int c = 0; c < C; c++) {
            e[c] = en[c];
        }
        if constexpr (!RED) {
            density_update();
        }

        /* record copy (solver_cpu_fdtd.tpp:288-325) */
        if ((rec_steps >> s) & 1ull) {
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
#pragma unroll
                for (int c = 0; c < C; c++) {
                    const int64_t g = gl + c;
                    if (g >= own_lo && g < own_hi &&
                        g >= (int64_t)R.gbegin &&
                        g < (int64_t)(R.gbegin + R.cols)) {
                        R.dst[row * R.cols + (uint64_t)(g - (int64_t)R.gbegin)] =
                            record_value<A>(k, R.type, R.row_idx, m[c], e[c],
                                            h[c], p[c], d[c]);
                    }
                }
            }
        }
    }

    /* write back the owned cells */
#pragma unroll
    for (int c = 0; c < C; c++) {
        const int64_t g = gl + c;
        if constexpr (A::SHIFT && std::is_same<ST, float>::value) {
            if (k.yshift) {
                add_shift<A>(k, m[c], d[c].v, -1.0);
            }
        }
        if (g >= own_lo && g < own_hi) {
            st<ST>(k.e_out, g, e[c]);
            st<ST>(k.h_out, g, h[c]);
            if constexpr (A::DC > 0) {
                st<ST>(k.p_out, g, p[c]);
            }
#pragma unroll
            for (int j = 0; j < A::DC; j++) {
                st<ST>(k.d_out, (uint64_t)j * k.dstride + g, d[c].v[j]);
            }
        }
    }
}

/* wavefronts per SIMD the compiler is asked to fit (1: no request): the
 * two-level Bloch kernel at four cells per lane fits three without spills
 * in its step loop, against two unasked */
#ifndef MBH_BLOCH2R_WAVES
#define MBH_BLOCH2R_WAVES 3
#endif
template<class A, int C>
constexpr int tile_min_waves()
{
    return (std::is_same<A, algo_bloch2r>::value && C == 4)
        ? MBH_BLOCH2R_WAVES
        : ((std::is_same<A, algo_bloch2>::value && C == 4) ? 3 : 1);
}

template<class A, int C, typename ST, bool RED = false>
__global__ __launch_bounds__(64)
__attribute__((amdgpu_waves_per_eu(tile_min_waves<A, C>()))) void
tile_kernel(KArgs k)
{
    tile_body<A, C, ST, RED>(k, k.tile_list ? k.tile_list[blockIdx.x]
                                            : blockIdx.x);
}

/* tile of W wavefronts (64 W threads) */
template<class A, int C, int W, typename ST>
__global__ __launch_bounds__(64 * W)
__attribute__((amdgpu_waves_per_eu(tile_min_waves<A, C>()))) void
tile_kernel_mw(KArgs k)
{
    tile_body<A, C, ST, false, W>(k, k.tile_list ? k.tile_list[blockIdx.x]
                                                 : blockIdx.x);
}

/* ---- launcher registry ------------------------------------------------- */
typedef hipError_t (*tile_launch_fn)(const KArgs& k, uint64_t num_tiles,
                                     hipStream_t stream);

struct kernel_entry {
    uint32_t algo;
    uint32_t levels;
    uint32_t form; /* MBH_FORM_* the kernel implements */
    uint32_t cells_per_lane;
    tile_launch_fn fn;
    uint32_t order; /* MBH_ORDER_* */
};

template<class A, int C, typename ST, bool RED = false>
hipError_t launch_tile(const KArgs& k, uint64_t num_tiles, hipStream_t stream)
{
    hipLaunchKernelGGL((tile_kernel<A, C, ST, RED>), dim3((unsigned)num_tiles),
                       dim3(64), 0, stream, k);
    return hipGetLastError();
}

/* multi-wavefront tiles: the tables list them with C W cells per lane of
 * the tile's first wavefront, so tile geometry (64 C W cells) needs no
 * other change on the host */
template<class A, int C, int W, typename ST>
hipError_t launch_tile_mw(const KArgs& k, uint64_t num_tiles,
                          hipStream_t stream)
{
    hipLaunchKernelGGL((tile_kernel_mw<A, C, W, ST>),
                       dim3((unsigned)num_tiles), dim3(64 * W), 0, stream, k);
    return hipGetLastError();
}

/* z-form state conversions (per cell, uniform material) */
struct ZArgs {
    const double* dref; /* reference coherence vectors, component-major */
    void* y;            /* z-form state (ST), component-major */
    uint64_t G;
    uint64_t dstride;
    const int32_t* mat;
    const double* zq;
    uint64_t zstride;
    uint32_t yshift; /* FP32 storage holds y - y* */
    uint64_t g0;     /* first super-grid cell converted */
    uint64_t count;  /* cells converted */
    double* out_cm;  /* y -> d only: if set, d of cell g0 + i goes to
                        out_cm[i * K + j] (cell-major) instead of dref */
};
typedef hipError_t (*zconv_fn)(const ZArgs& a, hipStream_t stream);

struct zconv_entry {
    uint32_t levels;
    uint32_t fp32;
    zconv_fn to_y;   /* dref -> y */
    zconv_fn from_y; /* y -> dref (written into dref) */
};

/* tables defined by the instantiation units */
extern const kernel_entry k_table_basic[];
extern const uint32_t k_table_basic_n;
extern const kernel_entry k_table_rodr[];
extern const uint32_t k_table_rodr_n;
extern const kernel_entry k_table_zform[];
extern const uint32_t k_table_zform_n;
extern const kernel_entry k_table_zm[];
extern const uint32_t k_table_zm_n;
extern const kernel_entry k_table_red[];
extern const uint32_t k_table_red_n;
extern const zconv_entry z_table[];
extern const uint32_t z_table_n;

} /* namespace mbh */

#endif
