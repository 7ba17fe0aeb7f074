/*
 * mbsolve_hip.h -- C ABI of the MI355X (gfx950) Maxwell-Bloch FDTD backend.
 *
 * This is the drop-in boundary below an mbsolve-style C++ solver class.
 * The caller (host C++ code that owns device/scenario/qm_description
 * objects) computes every FP64 constant on the host exactly as the
 * reference solver constructor does, packs them into the plain-old-data
 * structs below and hands them over.  No C++ or torch types cross this
 * boundary and no exception escapes it: every entry point returns an
 * integer status (MBH_OK == 0) and the message of the last failure is
 * available through mbh_last_error().
 *
 * Reference interfaces replaced (paths relative to mbsolve v0.4.0rc1):
 *   mbh_create        <- solver_cpu_fdtd<N,algo>::solver_cpu_fdtd
 *                        (solver-cpu/include/mbsolve/solver-cpu/
 *                         solver_cpu_fdtd.tpp:36-203): allocation :91-110,
 *                        per-cell initialisation :113-154, record scratch
 *                        :157-182, source table :184-202.
 *   mbh_run           <- solver_cpu_fdtd<N,algo>::run
 *                        (solver_cpu_fdtd.tpp:223-326): E update, sources,
 *                        H update, boundary condition, density update,
 *                        record copy to scratch.
 *   mbh_fetch_records <- bulk copy scratch -> result
 *                        (solver_cpu_fdtd.tpp:329-337).
 *   mbh_destroy       <- solver_cpu_fdtd<N,algo>::~solver_cpu_fdtd
 *                        (solver_cpu_fdtd.tpp:205-221).
 *   mbh_last_error    <- chk_err / std::runtime_error pattern
 *                        (solver-cuda/include/common_cuda.hpp:60-66).
 *   mbh_record        <- copy_list_entry_dev
 *                        (mbsolve-lib/include/mbsolve/lib/internal/
 *                         copy_list_entry.hpp:51-109).
 *   mbh_source        <- sim_source (internal/common_fdtd.hpp:39-45).
 *
 * Several independent runs (a parameter sweep) can be handed to one
 * mbh_create call; they are advanced together by the same kernels.  The
 * reference has one device/scenario per solver instance; with
 * num_runs == 1 the behaviour is exactly that of one reference solver.
 */
#ifndef MBSOLVE_HIP_H
#define MBSOLVE_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MBH_ABI_VERSION 1u

/* ---- status codes ------------------------------------------------------ */
enum {
    MBH_OK = 0,
    MBH_ERR_INVALID = 1,     /* invalid argument / inconsistent problem */
    MBH_ERR_HIP = 2,         /* HIP runtime error */
    MBH_ERR_NOMEM = 3,       /* device or host allocation failed */
    MBH_ERR_UNSUPPORTED = 4, /* valid, but not supported by this build */
    MBH_ERR_NODEVICE = 5     /* no usable GPU */
};

/* ---- density algorithms ------------------------------------------------ */
enum {
    MBH_ALGO_NOOP = 0,       /* lindblad_noop: pure Maxwell solver */
    MBH_ALGO_REG_CAYLEY = 1, /* lindblad_reg_cayley: density matrix,
                                Strang split, Cayley unitary step */
    MBH_ALGO_CVR_RODR = 2    /* lindblad_cvr_rodr: coherence vector,
                                generalized Rodrigues */
};

/* ---- numerical forms --------------------------------------------------- */
enum {
    /* FP64 state, FP64 arithmetic in the reference's operation order.
     * Results are bit-for-bit independent of tiling and temporal block. */
    MBH_FORM_FP64 = 0,
    /* FP32 state storage in HBM, FP64 arithmetic in registers.  Rodrigues
     * runs in the rotated basis (z-form, needs zform_consts); two-level
     * Cayley runs in closed Bloch form (the Cayley factor applied as a
     * rotation of the Bloch vector, constants derived from qm_consts);
     * three-level Cayley and noop keep the reference update. */
    MBH_FORM_MIXED = 1,
    /* FP64 state storage, same arithmetic as MIXED (z-form Rodrigues,
     * Bloch-form two-level Cayley) without the FP32 rounding of the stored
     * state. */
    MBH_FORM_ZFORM64 = 2,
    /* Rodrigues z-form, N = 3..6, FP32 state storage: rotation, affine part
     * and p in FP64, the generator product (W - I) r on the matrix cores as
     * FP16 hi + lo split products with FP32 accumulation. */
    MBH_FORM_MFMA = 3,
    /* as MBH_FORM_MFMA with FP64 state storage */
    MBH_FORM_MFMA64 = 4
};

/* ---- step orders ------------------------------------------------------- */
enum {
    /* solver_cpu_fdtd::run (solver_cpu_fdtd.tpp:230-325): E update,
     * sources, H update and boundary condition, density update, records */
    MBH_ORDER_PLAIN = 0,
    /* solver_cpu_fdtd_red::run (solver_cpu_fdtd_red.tpp:451-549): density
     * update, E update, sources, H update and boundary condition, records.
     * The density sees the field of the previous step, so inversion and
     * density records trail the plain order by one step.  FP64 form only. */
    MBH_ORDER_RED = 1
};

/* ---- record / source types (same values as mbsolve::record::type and
 *      mbsolve::source::type) --------------------------------------------- */
enum {
    MBH_REC_ELECTRIC = 0,
    MBH_REC_POLAR_DT = 1,
    MBH_REC_MAGNETIC = 2,
    MBH_REC_DENSITY = 3,
    MBH_REC_INVERSION = 4
};
enum { MBH_SRC_HARD = 0, MBH_SRC_SOFT = 1 };

/* Largest supported values. */
#define MBH_MAX_LEVELS 6u
#define MBH_MAX_SOURCES 4096u
#define MBH_MAX_RECORDS 4096u

/*
 * Per-material QM constant blob layouts (doubles), indexed by the offsets
 * below.  N = num_levels, K = N*N - 1, B = K / 2 (integer division).
 *
 * MBH_ALGO_REG_CAYLEY (lindblad_reg_cayley::sim_constants,
 * algo_lindblad_reg_cayley.hpp:59-87), complex matrices are N x N,
 * row-major, (re, im) interleaved; real matrices N x N row-major:
 *   [0] has_qm (1.0 / 0.0)       [1] carrier_dens
 *   [2] dt                       [3] prop_im  = (-1.0 / HBAR) * dt
 *   [4] rhs_im = -1.0 / HBAR
 *   [5 ...]  H0 (2N^2), mu (2N^2), d1 (N^2), d2 (N^2), D1 (N^2), D2 (N^2)
 *   stride = 5 + 8 N^2
 *
 * MBH_ALGO_CVR_RODR (lindblad_cvr_rodr::sim_constants,
 * algo_lindblad_cvr_rodr.hpp:55-96), K x K row-major:
 *   [0] has_qm  [1] M_CP  [2] d_t  [3] theta_1  [4] n_theta (count of
 *   non-zero rotation planes, <= B, sorted by descending theta)
 *   [5 ...]  M (K^2), U (K^2), U2 (K^2), A1 (K^2), v (K), d_eq (K),
 *            d_in (K), theta (B), C1 (B K^2), C2 (B K^2)
 *   stride = 5 + 4 K^2 + 3 K + B + 2 B K^2
 *
 * MBH_ALGO_NOOP: stride 0 (qm_consts may be NULL).
 *
 * Rodrigues z-form blob (zform_consts, required for MBH_FORM_MIXED and
 * MBH_FORM_ZFORM64 with MBH_ALGO_CVR_RODR).  With the real block
 * decomposition U = Q T Q^T (T = diag of [[0,-theta_j],[theta_j,0]]
 * blocks on coordinate pairs (2j, 2j+1), j < n_theta; zero elsewhere) and
 * y = Q^T (A1 (d + d_in) - d_in), one step at field e is
 *   r = R(e) y   (plane j rotated by phi_j = theta_dt_j * e)
 *   p = a . r + b
 *   y' = W r + c,  W = Q^T A1 A1 Q,  c = Q^T (A1 A1 - I) d_in
 *   d (state after the step, for records) = P r + o,
 *        P = A1 Q,  o = (A1 - I) d_in
 *   [0] has_qm  [1] n_theta  [2] b
 *   [3 ...]  theta_dt (B), W (K^2), c (K), a (K), P (K^2), o (K),
 *            Qt (K^2), QtA1 (K^2), d_in (K)
 *   stride = 3 + B + 4 K^2 + 4 K
 */
#define MBH_CAYLEY_STRIDE(N) (5u + 8u * (N) * (N))
#define MBH_RODR_K(N) ((N) * (N) - 1u)
#define MBH_RODR_STRIDE(N)                                                   \
    (5u + 4u * MBH_RODR_K(N) * MBH_RODR_K(N) + 3u * MBH_RODR_K(N) +          \
     MBH_RODR_K(N) / 2u +                                                    \
     2u * (MBH_RODR_K(N) / 2u) * MBH_RODR_K(N) * MBH_RODR_K(N))
#define MBH_ZFORM_STRIDE(N)                                                  \
    (3u + MBH_RODR_K(N) / 2u + 4u * MBH_RODR_K(N) * MBH_RODR_K(N) +          \
     4u * MBH_RODR_K(N))

/* Density state per cell (doubles):
 *   REG_CAYLEY: N x N complex matrix, row-major, (re, im) interleaved: 2N^2
 *   CVR_RODR:   coherence vector d, K = N^2 - 1
 *   NOOP:       0 */
#define MBH_DENSITY_STRIDE_CAYLEY(N) (2u * (N) * (N))
#define MBH_DENSITY_STRIDE_RODR(N) (MBH_RODR_K(N))

/* One source (sim_source): the values are source_data[k * num_timesteps + n]
 * for source k at step n (host-evaluated, source.hpp:94-106). */
typedef struct mbh_source {
    uint32_t type;  /* MBH_SRC_HARD: e[x] = value; MBH_SRC_SOFT: e[x] += */
    uint32_t x_idx; /* grid index, (unsigned)(position / dx) */
} mbh_source;

/* One record (copy_list_entry_dev).  At step n the record is written iff
 * floor(interval_ratio*(n-1)) != floor(interval_ratio*n) (n == 0 always),
 * to scratch[offset_real + floor(interval_ratio*n)*cols + (i - position_idx)]
 * for i in [position_idx, position_idx + cols).  Rows >= `rows` are
 * skipped (the reference writes them out of bounds). */
typedef struct mbh_record {
    uint32_t type;       /* MBH_REC_* */
    uint32_t row_idx;    /* density record: zero-based row */
    uint32_t col_idx;    /* density record: zero-based column */
    uint32_t is_complex; /* imaginary part is reserved but never written */
    uint64_t rows;
    uint64_t cols;
    uint64_t position_idx;
    double interval_ratio;
    uint64_t offset_real; /* in doubles, within this run's scratch */
    uint64_t offset_imag;
} mbh_record;

/* One simulation run: everything solver_cpu_fdtd's constructor derives
 * from (device, scenario).  All pointers are host pointers and are only
 * read during mbh_create. */
typedef struct mbh_problem {
    uint32_t abi_version;  /* MBH_ABI_VERSION */
    uint32_t algo;         /* MBH_ALGO_* */
    uint32_t num_levels;   /* N (0 for NOOP) */
    uint32_t num_materials;
    uint64_t num_gridpoints; /* N_x >= 1 */
    uint64_t num_timesteps;  /* n_t */
    double dx_inv;           /* 1 / gridpoint_size */
    const double* fdtd_consts; /* [num_materials][4] fac_a fac_b fac_c gamma */
    const double* qm_consts;   /* [num_materials][qm_stride] */
    uint64_t qm_stride;
    const double* zform_consts; /* [num_materials][zform_stride] or NULL */
    uint64_t zform_stride;
    const uint32_t* mat_idx; /* [num_gridpoints] */
    const double* e_init;    /* [num_gridpoints] */
    const double* h_init;    /* [num_gridpoints + 1]; h_init[0] and
                                h_init[num_gridpoints] are read by the first
                                step only (the boundary condition holds H
                                at zero there after every step) */
    const double* d_init;    /* [num_gridpoints][density stride] */
    uint32_t num_sources;
    const mbh_source* sources;
    const double* source_data; /* [num_sources][num_timesteps] */
    uint32_t num_records;
    const mbh_record* records;
    uint64_t scratch_size; /* doubles */
} mbh_problem;

typedef struct mbh_options {
    uint32_t form;           /* MBH_FORM_* */
    uint32_t temporal_block; /* time steps per kernel launch, 0 = auto;
                                at most 64 (a deeper request is clamped) */
    uint32_t tile_cells;     /* cells per wavefront tile incl. halo, 0 = auto */
    int32_t device;          /* HIP device ordinal, -1 = current */
    uint32_t graph_launches; /* kernel launches per captured hipGraph
                                (rounded up to an even count), 1 = no
                                graph, 0 = auto: graphs of about 2 ms of
                                launches when one launch of the tuned
                                geometry takes at most 0.25 ms and a full
                                run replays the graph at least twice, no
                                graph otherwise (mbh_info.graph_launches
                                reports the choice) */
    uint32_t step_order;     /* MBH_ORDER_* */
    uint32_t flags;          /* MBH_OPT_* */
    uint32_t reserved[1];
} mbh_options;

/* mbh_options.flags: run the general kernels of the form even where the
 * problem admits a specialised one (pure Maxwell with every held H cell
 * starting at zero; two-level Bloch form with a diagonal Hamiltonian, a
 * real off-diagonal dipole and lossless materials, held H starting at
 * zero).  The specialised kernels drop terms that are exact zeros,
 * so the results are the same; this flag is for verifying that. */
#define MBH_OPT_GENERIC 1u

typedef struct mbh_info {
    uint32_t form;
    uint32_t temporal_block;
    uint32_t tile_cells;   /* cells held by one wavefront (incl. halo) */
    uint32_t owned_cells;  /* cells written back per tile */
    uint64_t num_tiles;
    uint64_t total_cells;  /* sum over runs of num_gridpoints */
    uint64_t launches;     /* kernel launches issued by the last mbh_run */
    uint64_t state_bytes;  /* device bytes of one state buffer */
    uint32_t graph_launches;
    uint32_t step_order;   /* MBH_ORDER_* */
    uint32_t mfma_tiles;   /* MFMA forms: tiles whose density update runs on
                              the matrix cores (one material with a density) */
    uint32_t flags;        /* MBH_INFO_* */
} mbh_info;

/* mbh_info.flags: the pure-Maxwell FP32 kernel without the H factor of
 * the cells held at zero is in use (every such cell starts with H = 0) */
#define MBH_INFO_HZ_ZERO 1u
/* mbh_info.flags: the two-level Bloch kernel specialised for a diagonal
 * Hamiltonian, a real off-diagonal dipole and lossless materials, with
 * every held H cell starting at zero, is in use */
#define MBH_INFO_BLOCH_REAL 2u

typedef struct mbh_ctx mbh_ctx;

/* Library version string ("mbsolve-hip <abi> gfx950"). */
const char* mbh_version(void);

/* Number of visible HIP devices (0 when none). */
int mbh_device_count(int* count);

/* Default options: MBH_FORM_FP64, plain step order, automatic tiling,
 * current device. */
void mbh_default_options(mbh_options* opt);

/* Uploads constants, initial state, source tables and record descriptors
 * of num_runs runs (all with the same algo, num_levels and num_timesteps),
 * allocates the zeroed record scratch.  On success *out owns all device
 * memory and must be released with mbh_destroy. */
int mbh_create(const mbh_problem* runs,
               uint32_t num_runs,
               const mbh_options* opt,
               mbh_ctx** out);

/* Advances all runs through time steps n in [n_begin, n_end), in the step
 * order chosen at creation (solver_cpu_fdtd::run or solver_cpu_fdtd_red::run,
 * solver_cpu_fdtd_red.tpp:351-567).  Returns once the work is queued on the
 * context's stream; use mbh_synchronize or a fetch to wait. */
int mbh_run(mbh_ctx* ctx, uint64_t n_begin, uint64_t n_end);

/* Rewinds every run to the state handed to mbh_create (the state before
 * step 0) and zeroes the record scratch, so that mbh_run(ctx, 0, n_t)
 * repeats the simulation.  Queued on the context's stream.  No counterpart
 * in the reference, whose solver runs once per construction
 * (solver_cpu_fdtd.tpp:223-338): for repeated sweeps and benchmarks that
 * keep one context. */
int mbh_reset(mbh_ctx* ctx);

/* Waits for all queued work of the context. */
int mbh_synchronize(mbh_ctx* ctx);

/* Device time (HIP events on the context stream) of the last mbh_run. */
int mbh_elapsed_ms(mbh_ctx* ctx, double* ms);

/* Copies run `run`'s record scratch (scratch_size doubles) to dst. */
int mbh_fetch_records(mbh_ctx* ctx, uint32_t run, double* dst,
                      uint64_t count);

/* Copies run `run`'s current state to host arrays (any may be NULL):
 * e [N_x], h [N_x + 1], p [N_x], d [N_x][density stride]. */
int mbh_fetch_state(mbh_ctx* ctx, uint32_t run, double* e, double* h,
                    double* p, double* d);

/* Size of the per-GPU staging array of mbh_fetch_state_noalloc. */
#define MBH_NOALLOC_SCRATCH_BYTES (64u << 20)

/* As mbh_fetch_state, without allocating: no counterpart in the reference.
 * For callers that cannot allocate memory while they read the state (a
 * monitoring loop between mbh_run calls, harnesses that forbid hipMalloc).
 * mbh_fetch_state allocates a device buffer for the z-form read-out and
 * host buffers for FP32 widening and the density transposition; this entry
 * point instead stages every conversion (FP32 -> FP64, z-form or Bloch
 * state -> the reference density layout, component-major -> cell-major)
 * in one fixed device array of MBH_NOALLOC_SCRATCH_BYTES per GPU, held
 * from the first call on that GPU until the process ends, and copies from
 * there into the caller's host arrays.  The results are bit-identical to
 * mbh_fetch_state.  A run whose largest staged array (N_x + 1 doubles, or
 * N_x times the density stride when d is requested) exceeds the array is
 * rejected with MBH_ERR_INVALID.  Calls on one GPU share the array: calls
 * that overlap (several threads or contexts on one GPU at once) race. */
int mbh_fetch_state_noalloc(mbh_ctx* ctx, uint32_t run, double* e, double* h,
                            double* p, double* d);

/* Describes the chosen launch configuration. */
int mbh_get_info(const mbh_ctx* ctx, mbh_info* info);

/* Releases the context (NULL is ignored). */
void mbh_destroy(mbh_ctx* ctx);

/* Message of the last failure on ctx, or of the calling thread's last
 * failed call when ctx is NULL (e.g. a failed mbh_create). */
const char* mbh_last_error(const mbh_ctx* ctx);

#ifdef __cplusplus
}
#endif

#endif /* MBSOLVE_HIP_H */
