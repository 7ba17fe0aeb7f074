/*
 * C ABI of the gfx950 Maxwell-Bloch backend (declared in mbsolve_hip.h).
 *
 * Responsibilities: validate the problems, lay the runs out as one
 * super-grid, upload constants / state / source tables / record
 * descriptors, choose the tile geometry, and drive the time loop on one
 * HIP stream (optionally replaying a captured hipGraph of launches).
 */
#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include <mutex>

#include "mbh_zmfma.hpp"

using namespace mbh;

/*
 * Staging array of mbh_fetch_state_noalloc: one copy per GPU (placed when
 * this code object is loaded on that GPU, held until the process ends).
 * Host code never uses the name itself for device memory: the per-device
 * address comes from hipGetSymbolAddress (noalloc_scratch below).
 */
static __device__ __align__(256) unsigned char
    scratch[MBH_NOALLOC_SCRATCH_BYTES];

namespace {

thread_local std::string g_last_error;

/* small device helpers */
__global__ void set_u64(uint64_t* p, uint64_t v) { *p = v; }
__global__ void add_u64(uint64_t* p, uint64_t v) { *p += v; }

/* H of the trailing pseudo cell (h[N] of the last run) in both state
 * buffers, once step 0 has consumed it */
template<typename ST>
__global__ void zero_last_h(ST* h0, ST* h1, uint64_t g)
{
    h0[g] = ST(0);
    h1[g] = ST(0);
}

/* state read-out without allocation (mbh_fetch_state_noalloc) */
template<typename ST>
__global__ void widen_range(const ST* __restrict__ src,
                            double* __restrict__ dst, uint64_t n)
{
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += stride) {
        dst[i] = (double)src[i];
    }
}

/* component-major density [K][dstride] of cells [base, base + cells) ->
 * cell-major [cells][K] */
template<typename ST>
__global__ void gather_cell_major(const ST* __restrict__ d, uint64_t dstride,
                                  uint64_t base, uint64_t cells, uint32_t K,
                                  double* __restrict__ out)
{
    const uint64_t n = cells * K;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n;
         t += stride) {
        const uint64_t j = t / cells, i = t - j * cells;
        out[i * K + j] = (double)d[j * dstride + base + i];
    }
}

/* carried Bloch state -> reference 2x2 complex layout (the device twin of
 * bloch2_from_sigma plus the Hermitian fill of mbh_fetch_state) */
template<typename ST>
__global__ void bloch_to_ref(const ST* __restrict__ sig, uint64_t dstride,
                             uint64_t base, uint64_t cells,
                             const int32_t* __restrict__ mat,
                             const double* __restrict__ qm, uint64_t qstride,
                             double* __restrict__ out)
{
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
         i < cells; i += stride) {
        const uint64_t g = base + i;
        const double* o = qm + (uint64_t)mat[g] * qstride;
        double sv[4], rho[4];
        for (int j = 0; j < 4; j++) {
            sv[j] = (double)sig[(uint64_t)j * dstride + g];
        }
        if (o[0] == 0.0) {
            for (int j = 0; j < 4; j++) {
                rho[j] = sv[j];
            }
        } else {
            const double* R = o + 18;
            const double gg = o[26];
            rho[0] = R[0] * sv[0] + R[1] * sv[3];
            rho[1] = R[2] * sv[0] + R[3] * sv[3];
            rho[2] = gg * sv[1];
            rho[3] = -gg * sv[2];
        }
        double* di = out + i * 8;
        di[0] = rho[0];
        di[1] = 0.0;
        di[2] = rho[2];
        di[3] = rho[3];
        di[4] = rho[2];
        di[5] = -rho[3];
        di[6] = rho[1];
        di[7] = 0.0;
    }
}

__global__ void fill_zero_f64(double* p, uint64_t n)
{
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += stride) {
        p[i] = 0.0;
    }
}

struct dev_buf {
    void* p = nullptr;
    size_t bytes = 0;
};

struct run_layout {
    uint64_t base;      /* super-grid cell of cell 0 */
    uint64_t cells;     /* N_x */
    uint64_t scratch;   /* doubles */
    uint64_t scratch_off; /* doubles, into the shared scratch allocation */
    uint32_t dstride_ref; /* reference density stride */
};

/* Components of the reference 2x2 density layout (re/im interleaved,
 * row-major) kept by the Bloch form: rho00, rho11, Re rho01, Im rho01. */
constexpr uint32_t BLOCH_FROM_REF[4] = { 0, 6, 2, 3 };

/* Bloch-form constants (layout in algo_bloch2) from the reference Cayley
 * blob of a two-level material (MBH_CAYLEY_STRIDE(2) doubles). */
void bloch2_constants(const double* cq, double* o)
{
    typedef std::complex<double> cx;
    std::fill(o, o + BLOCH2_STRIDE, 0.0);
    if (cq[0] == 0.0) {
        return;
    }
    const double carrier = cq[1], prop_im = cq[3], rhs_im = cq[4];
    cx H0[4], mu[4];
    for (int i = 0; i < 4; i++) {
        H0[i] = cx(cq[5 + 2 * i], cq[6 + 2 * i]);
        mu[i] = cx(cq[13 + 2 * i], cq[14 + 2 * i]);
    }
    const double* r1 = cq + 21;
    const double* r2 = cq + 25;
    const double* D1 = cq + 29;
    const double* D2 = cq + 33;
    const double g = D2[1];
    /* Pauli components of a Hermitian 2x2: X = x0 I + x.sigma */
    auto pauli = [](const cx* X, double* v) {
        v[0] = 0.5 * (X[0].real() + X[3].real());
        v[1] = X[1].real();
        v[2] = -X[1].imag();
        v[3] = 0.5 * (X[0].real() - X[3].real());
    };
    double h[4], m[4];
    pauli(H0, h);
    pauli(mu, m);
    /* H = H0 - e mu; beta = -prop_im h_vec, alpha0 = -(prop_im / 2) h0 */
    o[0] = 1.0;
    for (int j = 0; j < 3; j++) {
        o[1 + j] = -prop_im * h[1 + j];
        o[4 + j] = prop_im * m[1 + j];
    }
    o[7] = -0.5 * prop_im * h[0];
    o[8] = 0.5 * prop_im * m[0];
    o[9] = g * g;
    /* (a, b) = Ti (tau, m_z), (tau, m_z) = T (a, b) */
    const double T[4] = { 1.0, 1.0, 0.5, -0.5 };
    const double Ti[4] = { 0.5, 1.0, 0.5, -1.0 };
    auto mul2 = [](const double* A, const double* B, double* C) {
        C[0] = A[0] * B[0] + A[1] * B[2];
        C[1] = A[0] * B[1] + A[1] * B[3];
        C[2] = A[2] * B[0] + A[3] * B[2];
        C[3] = A[2] * B[1] + A[3] * B[3];
    };
    double E[4], ETi[4], R[4];
    mul2(D1, D1, E);
    mul2(E, Ti, ETi);
    mul2(T, ETi, o + 10); /* F = T D1^2 Ti */
    mul2(D1, Ti, R);      /* populations after the step */
    /* p = carrier Re tr[(-i/hbar [H0, d] + D(d)) mu] as a linear form of
     * (rho00, rho11, Re rho01, Im rho01), evaluated on basis states */
    const cx I(0.0, 1.0);
    double pc[4];
    for (int b = 0; b < 4; b++) {
        cx d[4] = { 0.0, 0.0, 0.0, 0.0 };
        if (b == 0) {
            d[0] = 1.0;
        } else if (b == 1) {
            d[3] = 1.0;
        } else if (b == 2) {
            d[1] = 1.0;
            d[2] = 1.0;
        } else {
            d[1] = I;
            d[2] = -I;
        }
        cx rhs[4];
        for (int i = 0; i < 2; i++) {
            for (int k = 0; k < 2; k++) {
                cx c = 0.0;
                for (int j = 0; j < 2; j++) {
                    c += H0[i * 2 + j] * d[j * 2 + k] -
                        d[i * 2 + j] * H0[j * 2 + k];
                }
                cx dis = (i == k) ? r1[i * 2] * d[0] + r1[i * 2 + 1] * d[3]
                                  : r2[i * 2 + k] * d[i * 2 + k];
                rhs[i * 2 + k] = I * rhs_im * c + dis;
            }
        }
        double tr = 0.0;
        for (int i = 0; i < 2; i++) {
            for (int k = 0; k < 2; k++) {
                tr += (rhs[i * 2 + k] * mu[k * 2 + i]).real();
            }
        }
        pc[b] = carrier * tr;
    }
    /* rho_n = D(sigma): (rho00, rho11) = R (tau, m_z), rho01 = g (m_x - i m_y) */
    o[14] = pc[0] * R[0] + pc[1] * R[2];
    o[15] = pc[2] * g;
    o[16] = -pc[3] * g;
    o[17] = pc[0] * R[1] + pc[1] * R[3];
    for (int j = 0; j < 4; j++) {
        o[18 + j] = R[j];
        o[22 + j] = D1[j];
    }
    o[26] = g;
    o[27] = std::fma(o[7], o[7], 1.0); /* 1 + alpha0^2 at e = 0 */
}

/* reference (rho00, rho11, Re rho01, Im rho01) -> carried Bloch state
 * sigma with D(sigma) = rho */
void bloch2_to_sigma(const double* o, const double* rho, double* sig)
{
    if (o[0] == 0.0) {
        std::copy(rho, rho + 4, sig);
        return;
    }
    const double* D1 = o + 22;
    const double g = o[26];
    const double det = D1[0] * D1[3] - D1[1] * D1[2];
    const double a = (D1[3] * rho[0] - D1[1] * rho[1]) / det;
    const double b = (D1[0] * rho[1] - D1[2] * rho[0]) / det;
    sig[0] = a + b;
    sig[1] = rho[2] / g;
    sig[2] = -rho[3] / g;
    sig[3] = 0.5 * (a - b);
}

/* carried Bloch state -> reference (rho00, rho11, Re rho01, Im rho01) */
void bloch2_from_sigma(const double* o, const double* sig, double* rho)
{
    if (o[0] == 0.0) {
        std::copy(sig, sig + 4, rho);
        return;
    }
    const double* R = o + 18;
    const double g = o[26];
    rho[0] = R[0] * sig[0] + R[1] * sig[3];
    rho[1] = R[2] * sig[0] + R[3] * sig[3];
    rho[2] = g * sig[1];
    rho[3] = -g * sig[2];
}

/* y* = -Q^T d_in of a z-form blob (layout algo_zform<N>): the state of
 * the relaxation equilibrium d = -d_in, the fixed point of the step at
 * zero field */
void zform_yeq(uint32_t levels, const double* zq, double* yeq)
{
    const int K = (int)(levels * levels) - 1;
    const int B = K / 2;
    const double* Qt = zq + 3 + B + 2 * K * K + 3 * K; /* OFF_QT */
    const double* din = Qt + 2 * K * K;                /* OFF_DIN */
    for (int i = 0; i < K; i++) {
        double acc = 0.0;
        for (int j = 0; j < K; j++) {
            acc += Qt[i * K + j] * din[j];
        }
        yeq[i] = (zq[0] != 0.0) ? -acc : 0.0;
    }
}

/* MFMA-form constants of one material (layout zm_geom<N>) from its
 * z-form blob (layout algo_zform<N>): split FP16 fragments of the scaled
 * W - I in the kernel's row order, and per-lane rows of theta, a, c and
 * the diagonal rows of P, o. */
template<int N>
void zm_build(const double* zq, unsigned char* out)
{
    using Gm = zm_geom<N>;
    using Z = algo_zform<N>;
    constexpr int K = Gm::K;
    std::memset(out, 0, Gm::BYTES);
    double* hd = reinterpret_cast<double*>(out);
    if (zq[0] == 0.0) {
        return;
    }
    static_assert(Z::OFF_QT == 3 + Gm::B + 2 * K * K + 3 * K &&
                      Z::OFF_DIN == Z::OFF_QT + 2 * K * K,
                  "z-form blob layout (zform_yeq)");
    const double* W = zq + Z::OFF_W;
    const double* cv = zq + Z::OFF_C;
    double dmax = 0.0;
    for (int i = 0; i < K; i++) {
        for (int k = 0; k < K; k++) {
            dmax = std::max(dmax, std::fabs(W[i * K + k] - (i == k ? 1.0 : 0.0)));
        }
        if (Gm::CFOLD) {
            dmax = std::max(dmax, std::fabs(cv[i]));
        }
    }
    /* scale 2^s puts max |W - I| near 2^13 (FP16 holds up to 65504) */
    const int s = (dmax > 0.0) ? 13 - (int)std::ceil(std::log2(dmax)) : 0;
    const double sc = std::ldexp(1.0, s);
    double thmax = 0.0;
    for (int b = 0; b < Gm::B; b++) {
        thmax = std::max(thmax, std::fabs(zq[Z::OFF_TH + b]));
    }
    hd[0] = std::ldexp(1.0, -s);
    hd[1] = thmax;
    hd[2] = zq[2];
    hd[3] = 1.0;
    _Float16* fr = reinterpret_cast<_Float16*>(out + Gm::FRAG);
    for (int t = 0; t < Gm::MT; t++) {
        for (int kc = 0; kc < Gm::KC; kc++) {
            for (int l = 0; l < 64; l++) {
                const int i = Gm::out_row(t, l % 16), qq = l / 16;
                for (int sl = 0; sl < 8; sl++) {
                    const int u = 8 * kc + sl;
                    const int col = (u < Gm::ROWS) ? Gm::slot_row(qq, u) : K;
                    double v = 0.0;
                    if (i < K && col < K) {
                        v = (W[i * K + col] - (i == col ? 1.0 : 0.0)) * sc;
                    } else if (i < K && Gm::CFOLD && u == Gm::ROWS &&
                               qq == 0) {
                        v = cv[i] * sc;
                    }
                    const _Float16 hi = (_Float16)v;
                    const double r1 = v - (double)hi;
                    const _Float16 mid = (_Float16)r1;
                    const _Float16 lo = (_Float16)(r1 - (double)mid);
                    const int f = (t * Gm::KC + kc) * Gm::AP;
                    fr[(f * 64 + l) * 8 + sl] = hi;
                    if (Gm::AP == 3) {
                        fr[((f + 1) * 64 + l) * 8 + sl] = mid;
                        fr[((f + 2) * 64 + l) * 8 + sl] = lo;
                    } else {
                        fr[((f + 1) * 64 + l) * 8 + sl] = (_Float16)r1;
                    }
                }
            }
        }
    }
    double* th = reinterpret_cast<double*>(out + Gm::TH);
    double* ar = reinterpret_cast<double*>(out + Gm::AR);
    double* cr = reinterpret_cast<double*>(out + Gm::CR);
    double* pr = reinterpret_cast<double*>(out + Gm::PR);
    double* ov = reinterpret_cast<double*>(out + Gm::OV);
    for (int l = 0; l < 64; l++) {
        const int qq = l / 16;
        for (int i = 0; i < Gm::PL; i++) {
            const int r0 = Gm::slot_row(qq, 2 * i);
            const int plane = r0 / 2;
            th[i * 64 + l] =
                (r0 < K && plane < Gm::B) ? zq[Z::OFF_TH + plane] : 0.0;
        }
        for (int i = 0; i < Gm::ROWS; i++) {
            const int row = Gm::slot_row(qq, i);
            const bool ok = row < K;
            ar[i * 64 + l] = ok ? zq[Z::OFF_A + row] : 0.0;
            cr[i * 64 + l] = ok ? zq[Z::OFF_C + row] : 0.0;
            for (int kk = 0; kk < Gm::NP; kk++) {
                const int d = N * (N - 1) + kk;
                pr[(kk * Gm::ROWS + i) * 64 + l] =
                    ok ? zq[Z::OFF_P + d * K + row] : 0.0;
            }
        }
    }
    for (int kk = 0; kk < Gm::NP; kk++) {
        ov[kk] = zq[Z::OFF_O + N * (N - 1) + kk];
    }
}

uint64_t zm_bytes(uint32_t levels)
{
    switch (levels) {
    case 3: return zm_geom<3>::BYTES;
    case 4: return zm_geom<4>::BYTES;
    case 5: return zm_geom<5>::BYTES;
    case 6: return zm_geom<6>::BYTES;
    default: return 0;
    }
}

void zm_build_levels(uint32_t levels, const double* zq, unsigned char* out)
{
    switch (levels) {
    case 3: zm_build<3>(zq, out); break;
    case 4: zm_build<4>(zq, out); break;
    case 5: zm_build<5>(zq, out); break;
    case 6: zm_build<6>(zq, out); break;
    default: break;
    }
}

} /* namespace */

struct mbh_ctx {
    std::string err;
    int device = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    bool timed = false;

    uint32_t algo = 0, levels = 0, form = 0;
    uint32_t order = MBH_ORDER_PLAIN;
    uint64_t n_t = 0;
    uint32_t DC = 0;        /* stored density components per cell */
    uint32_t dref = 0;      /* reference density components per cell */
    bool fp32 = false;
    bool zform = false;
    bool bloch = false; /* two-level Cayley in closed Bloch form */
    /* every Bloch material has beta_y = 0, a field-free alpha0, no p
     * term in tau and m_z, and no loss (with hz_zero, algo_bloch2r
     * applies) */
    bool bloch_real = false;
    bool generic = false; /* MBH_OPT_GENERIC: no specialised kernels */
    std::vector<int32_t> mat_h; /* per super-grid cell (Bloch conversions) */
    std::vector<double> qm_h;   /* device blob copy (Bloch conversions) */

    uint64_t G = 0, dstride = 0;
    std::vector<run_layout> runs;

    /* tiling */
    uint32_t C = 1, halo = 1, owned = 62, tb = 1;
    uint64_t num_tiles = 0;
    tile_launch_fn launch = nullptr;
    zconv_fn to_y = nullptr, from_y = nullptr;

    /* device memory */
    std::vector<void*> allocs;
    void* state[2][4] = { { nullptr } }; /* [buffer][e,h,p,d] */
    void* init[4] = { nullptr };         /* initial state (mbh_reset) */
    uint64_t state_elems[4] = { 0 };     /* elements per state array */
    int cur = 0;
    int32_t* mat = nullptr;
    uint8_t* hzero = nullptr;
    double* fdtd = nullptr;
    double* qm = nullptr;
    uint64_t qstride = 0;
    void* frag = nullptr; /* MFMA form constants */
    uint64_t frag_stride = 0;
    /* MFMA form: tiles of one material run on the matrix cores, the others
     * the VALU z-form body in the same kernel; tiles_a lists the mixed
     * tiles first so that their longer waves start early */
    bool split = false;
    /* every cell with H held at zero starts with H = 0 */
    bool hz_zero = false;
    uint32_t* tiles_a = nullptr;
    int32_t* tiles_mat = nullptr; /* per list position: material | QM flag */
    uint64_t n_tiles_mixed = 0;
    uint64_t n_tiles_mfma = 0;      /* uniform tiles of a density material */
    std::vector<uint8_t> mat_has_qm; /* per material (MFMA forms) */
    dev_source* src = nullptr;
    uint32_t* tile_src = nullptr;
    dev_record* rec = nullptr;
    uint32_t* tile_rec_ptr = nullptr;
    uint32_t* tile_rec_idx = nullptr;
    double* src_data = nullptr;
    double* rec_buf = nullptr; /* record scratch of all runs */
    uint64_t rec_total = 0;    /* doubles in rec_buf */
    uint64_t* n_base = nullptr;

    /* host copies for rebuilding per-geometry tile tables */
    std::vector<dev_source> srcs;
    std::vector<dev_record> recs;
    uint64_t tuned_candidates = 0;

    /* graphs */
    uint32_t graph_launches = 1;
    hipGraphExec_t graph[2] = { nullptr, nullptr }; /* per start buffer */
    double launch_ms_est = 0.0; /* one launch of the chosen geometry */
    uint32_t kform = 0;         /* kernel table form in use */

    uint64_t last_launches = 0;
    uint64_t state_bytes = 0;

    ~mbh_ctx()
    {
        if (device >= 0) {
            (void)hipSetDevice(device);
        }
        for (auto& g : graph) {
            if (g) {
                (void)hipGraphExecDestroy(g);
            }
        }
        if (stream) {
            (void)hipStreamSynchronize(stream);
        }
        for (void* p : allocs) {
            (void)hipFree(p);
        }
        if (ev0) {
            (void)hipEventDestroy(ev0);
        }
        if (ev1) {
            (void)hipEventDestroy(ev1);
        }
        if (stream) {
            (void)hipStreamDestroy(stream);
        }
    }
};

namespace {

int fail(mbh_ctx* ctx, int code, const char* fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_last_error = buf;
    if (ctx) {
        ctx->err = buf;
    }
    return code;
}

/* A failed HIP call also leaves its code as the thread's last error,
 * which a later launch check (hipGetLastError after a launch) would
 * report as its own: the failure paths read it off. */
#define HIP_TRY(ctx, call)                                                    \
    do {                                                                      \
        hipError_t e_ = (call);                                               \
        if (e_ != hipSuccess) {                                               \
            (void)hipGetLastError();                                          \
            return fail(ctx, (e_ == hipErrorOutOfMemory) ? MBH_ERR_NOMEM       \
                                                         : MBH_ERR_HIP,       \
                        "%s failed: %s", #call, hipGetErrorString(e_));       \
        }                                                                     \
    } while (0)

int dalloc(mbh_ctx* ctx, void** p, size_t bytes)
{
    *p = nullptr;
    if (bytes == 0) {
        bytes = 8;
    }
    hipError_t e = hipMalloc(p, bytes);
    if (e != hipSuccess) {
        *p = nullptr;
        (void)hipGetLastError(); /* not to be reported by a later launch */
        return fail(ctx, (e == hipErrorOutOfMemory) ? MBH_ERR_NOMEM
                                                    : MBH_ERR_HIP,
                    "hipMalloc(%zu bytes) failed: %s", bytes,
                    hipGetErrorString(e));
    }
    ctx->allocs.push_back(*p);
    return MBH_OK;
}

/* Device address of `scratch` on the context's GPU (current), looked up
 * once per GPU and kept for the process. */
constexpr int NOALLOC_MAX_DEVICES = 64;
std::mutex g_noalloc_mu;
void* g_noalloc_ptr[NOALLOC_MAX_DEVICES] = { nullptr };

int noalloc_scratch(mbh_ctx* ctx, double** out)
{
    std::lock_guard<std::mutex> lock(g_noalloc_mu);
    if (ctx->device < 0 || ctx->device >= NOALLOC_MAX_DEVICES) {
        return fail(ctx, MBH_ERR_UNSUPPORTED,
                    "no staging array for device %d", ctx->device);
    }
    if (!g_noalloc_ptr[ctx->device]) {
        void* p = nullptr;
        hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            return fail(ctx, MBH_ERR_HIP,
                        "hipGetSymbolAddress(scratch) on device %d: %s",
                        ctx->device, hipGetErrorString(e));
        }
        g_noalloc_ptr[ctx->device] = p;
    }
    *out = static_cast<double*>(g_noalloc_ptr[ctx->device]);
    return MBH_OK;
}

template<typename T>
int upload(mbh_ctx* ctx, T** dst, const T* src, size_t count)
{
    int rc = dalloc(ctx, (void**)dst, count * sizeof(T));
    if (rc) {
        return rc;
    }
    if (count) {
        HIP_TRY(ctx, hipMemcpy(*dst, src, count * sizeof(T),
                               hipMemcpyHostToDevice));
    }
    return MBH_OK;
}

uint32_t ref_density_stride(uint32_t algo, uint32_t N)
{
    switch (algo) {
    case MBH_ALGO_REG_CAYLEY:
        return MBH_DENSITY_STRIDE_CAYLEY(N);
    case MBH_ALGO_CVR_RODR:
        return MBH_DENSITY_STRIDE_RODR(N);
    default:
        return 0;
    }
}

uint64_t qm_stride_of(uint32_t algo, uint32_t N)
{
    switch (algo) {
    case MBH_ALGO_REG_CAYLEY:
        return MBH_CAYLEY_STRIDE(N);
    case MBH_ALGO_CVR_RODR:
        return MBH_RODR_STRIDE(N);
    default:
        return 0;
    }
}

int validate(const mbh_problem* p, uint32_t r, const mbh_problem* p0,
             uint32_t form)
{
    if (p->abi_version != MBH_ABI_VERSION) {
        return fail(nullptr, MBH_ERR_INVALID,
                    "run %u: abi_version %u, library expects %u", r,
                    p->abi_version, MBH_ABI_VERSION);
    }
    if (p->algo > MBH_ALGO_CVR_RODR) {
        return fail(nullptr, MBH_ERR_INVALID, "run %u: unknown algo %u", r,
                    p->algo);
    }
    if (p->algo != MBH_ALGO_NOOP &&
        (p->num_levels < 2 || p->num_levels > MBH_MAX_LEVELS)) {
        return fail(nullptr, MBH_ERR_INVALID,
                    "run %u: num_levels %u outside [2, %u]", r, p->num_levels,
                    MBH_MAX_LEVELS);
    }
    if (p->algo != p0->algo || p->num_levels != p0->num_levels ||
        p->num_timesteps != p0->num_timesteps) {
        return fail(nullptr, MBH_ERR_INVALID,
                    "run %u: algo, num_levels and num_timesteps must be equal "
                    "for all runs of one context",
                    r);
    }
    if (p->num_gridpoints < 1 || p->num_gridpoints > (1ull << 40)) {
        return fail(nullptr, MBH_ERR_INVALID, "run %u: bad num_gridpoints",
                    r);
    }
    if (p->num_materials < 1 || !p->fdtd_consts || !p->mat_idx ||
        !p->e_init || !p->h_init) {
        return fail(nullptr, MBH_ERR_INVALID,
                    "run %u: missing materials or initial state", r);
    }
    const uint64_t qs = qm_stride_of(p->algo, p->num_levels);
    if (p->algo != MBH_ALGO_NOOP) {
        if (p->qm_stride != qs || !p->qm_consts || !p->d_init) {
            return fail(nullptr, MBH_ERR_INVALID,
                        "run %u: qm_stride %llu (expected %llu) or missing "
                        "qm_consts/d_init",
                        r, (unsigned long long)p->qm_stride,
                        (unsigned long long)qs);
        }
    }
    const bool mfma = (form == MBH_FORM_MFMA || form == MBH_FORM_MFMA64);
    if (mfma && (p->algo != MBH_ALGO_CVR_RODR || p->num_levels < 3)) {
        return fail(nullptr, MBH_ERR_UNSUPPORTED,
                    "run %u: the MFMA form is for Rodrigues with 3 to 6 "
                    "levels",
                    r);
    }
    if (p->algo == MBH_ALGO_CVR_RODR &&
        (form == MBH_FORM_MIXED || form == MBH_FORM_ZFORM64 || mfma)) {
        if (!p->zform_consts ||
            p->zform_stride != MBH_ZFORM_STRIDE(p->num_levels)) {
            return fail(nullptr, MBH_ERR_INVALID,
                        "run %u: the rotated-basis form needs zform_consts "
                        "with stride %u",
                        r, MBH_ZFORM_STRIDE(p->num_levels));
        }
    }
    for (uint64_t i = 0; i < p->num_gridpoints; i++) {
        if (p->mat_idx[i] >= p->num_materials) {
            return fail(nullptr, MBH_ERR_INVALID,
                        "run %u: mat_idx[%llu] = %u >= num_materials", r,
                        (unsigned long long)i, p->mat_idx[i]);
        }
    }
    if (p->num_sources > MBH_MAX_SOURCES ||
        (p->num_sources && (!p->sources || !p->source_data))) {
        return fail(nullptr, MBH_ERR_INVALID, "run %u: bad source list", r);
    }
    for (uint32_t k = 0; k < p->num_sources; k++) {
        if (p->sources[k].x_idx >= p->num_gridpoints ||
            p->sources[k].type > MBH_SRC_SOFT) {
            return fail(nullptr, MBH_ERR_INVALID,
                        "run %u: source %u at x_idx %u outside the grid or of "
                        "unknown type",
                        r, k, p->sources[k].x_idx);
        }
    }
    if (p->num_records > MBH_MAX_RECORDS ||
        (p->num_records && !p->records)) {
        return fail(nullptr, MBH_ERR_INVALID, "run %u: bad record list", r);
    }
    for (uint32_t k = 0; k < p->num_records; k++) {
        const mbh_record& R = p->records[k];
        if (R.type > MBH_REC_INVERSION || R.cols < 1 ||
            R.position_idx + R.cols > p->num_gridpoints ||
            R.offset_real + R.rows * R.cols > p->scratch_size ||
            !(R.interval_ratio > 0.0) || R.interval_ratio > 1.0) {
            return fail(nullptr, MBH_ERR_INVALID,
                        "run %u: record %u is inconsistent (type %u, position "
                        "%llu, cols %llu, rows %llu)",
                        r, k, R.type, (unsigned long long)R.position_idx,
                        (unsigned long long)R.cols,
                        (unsigned long long)R.rows);
        }
        if (R.type == MBH_REC_DENSITY && p->algo != MBH_ALGO_NOOP &&
            (R.row_idx >= p->num_levels || R.col_idx >= p->num_levels)) {
            return fail(nullptr, MBH_ERR_INVALID,
                        "run %u: density record %u index out of range", r, k);
        }
    }
    return MBH_OK;
}

const kernel_entry* find_kernel(uint32_t algo, uint32_t levels, uint32_t form,
                                uint32_t C, uint32_t order)
{
    const kernel_entry* tabs[5] = { k_table_basic, k_table_rodr,
                                    k_table_zform, k_table_zm, k_table_red };
    const uint32_t ns[5] = { k_table_basic_n, k_table_rodr_n,
                             k_table_zform_n, k_table_zm_n, k_table_red_n };
    for (int t = 0; t < 5; t++) {
        for (uint32_t i = 0; i < ns[t]; i++) {
            const kernel_entry& k = tabs[t][i];
            if (k.algo == algo && k.levels == levels && k.form == form &&
                k.cells_per_lane == C && k.order == order) {
                return &k;
            }
        }
    }
    return nullptr;
}

/* Kernel form implementing the requested numerical form: Rodrigues and
 * two-level Cayley have restructured kernels for MIXED and ZFORM64; the
 * others fall back to the reference operation order (ZFORM64 -> FP64). */
uint32_t kernel_form(uint32_t algo, uint32_t levels, uint32_t form)
{
    const bool restructured = (algo == MBH_ALGO_CVR_RODR) ||
        (algo == MBH_ALGO_REG_CAYLEY && levels == 2);
    if (!restructured && form == MBH_FORM_ZFORM64) {
        return MBH_FORM_FP64;
    }
    return form;
}

/* The z-forms with FP32 storage (VALU and MFMA) keep y - y* in memory (y*
 * appended to the device blobs), so the storage rounding scales with the
 * distance from the relaxation equilibrium. */
uint32_t yshift(const mbh_ctx* ctx)
{
    return (ctx->zform && ctx->fp32) ? 1u : 0u;
}

/* Cells owned by tiles: the trailing pseudo cell is never written. */
uint64_t owned_cells(const mbh_ctx* ctx) { return ctx->G - 1; }

void set_geometry(mbh_ctx* ctx, uint32_t C, uint32_t tb)
{
    ctx->C = C;
    ctx->tb = tb;
    ctx->halo = tb;
    ctx->owned = 64u * C - 2u * tb;
    ctx->num_tiles = (owned_cells(ctx) + ctx->owned - 1) / ctx->owned;
}

/* Per-tile source ranges (sources inside the loaded span, halo included)
 * and record lists (records overlapping the owned span). */
int build_tile_tables(mbh_ctx* ctx, bool with_records)
{
    const uint64_t nt = ctx->num_tiles;
    std::vector<uint32_t> ts(2 * nt);
    for (uint64_t t = 0; t < nt; t++) {
        const int64_t lo = (int64_t)(t * ctx->owned) - ctx->halo;
        const int64_t hi = lo + 64 * (int64_t)ctx->C;
        auto a = std::lower_bound(ctx->srcs.begin(), ctx->srcs.end(), lo,
                                  [](const dev_source& s, int64_t v) {
                                      return (int64_t)s.gx < v;
                                  });
        auto b = std::lower_bound(ctx->srcs.begin(), ctx->srcs.end(), hi,
                                  [](const dev_source& s, int64_t v) {
                                      return (int64_t)s.gx < v;
                                  });
        ts[2 * t] = (uint32_t)(a - ctx->srcs.begin());
        ts[2 * t + 1] = (uint32_t)(b - ctx->srcs.begin());
    }
    std::vector<uint32_t> ptr(nt + 1, 0), idx;
    if (with_records) {
        std::vector<std::pair<uint64_t, uint32_t>> tile_rec;
        for (uint32_t id = 0; id < ctx->recs.size(); id++) {
            const dev_record& d = ctx->recs[id];
            const uint64_t t0 = d.gbegin / ctx->owned;
            const uint64_t t1 = (d.gbegin + d.cols - 1) / ctx->owned;
            for (uint64_t t = t0; t <= t1 && t < nt; t++) {
                tile_rec.emplace_back(t, id);
            }
        }
        std::stable_sort(tile_rec.begin(), tile_rec.end(),
                         [](const std::pair<uint64_t, uint32_t>& x,
                            const std::pair<uint64_t, uint32_t>& y) {
                             return x.first < y.first;
                         });
        idx.reserve(tile_rec.size());
        for (auto& tr : tile_rec) {
            ptr[tr.first + 1]++;
            idx.push_back(tr.second);
        }
        for (uint64_t t = 0; t < nt; t++) {
            ptr[t + 1] += ptr[t];
        }
    }
    /* replace the previous tables */
    auto release = [ctx](void* p) {
        if (!p) {
            return;
        }
        auto it = std::find(ctx->allocs.begin(), ctx->allocs.end(), p);
        if (it != ctx->allocs.end()) {
            ctx->allocs.erase(it);
        }
        (void)hipFree(p);
    };
    release(ctx->tile_src);
    release(ctx->tile_rec_ptr);
    release(ctx->tile_rec_idx);
    release(ctx->tiles_a);
    release(ctx->tiles_mat);
    ctx->tile_src = ctx->tile_rec_ptr = ctx->tile_rec_idx = nullptr;
    ctx->tiles_a = nullptr;
    ctx->tiles_mat = nullptr;
    ctx->n_tiles_mixed = 0;
    ctx->n_tiles_mfma = 0;
    int rc;
    if (ctx->split) {
        /* a tile is uniform when all its 64 cells lie in the grid and have
         * the same material */
        std::vector<uint32_t> la, lb;
        for (uint64_t t = 0; t < nt; t++) {
            const int64_t lo = (int64_t)(t * ctx->owned) - ctx->halo;
            const int64_t hi = lo + 64 * (int64_t)ctx->C;
            bool uni = lo >= 0 && hi <= (int64_t)ctx->G;
            if (uni) {
                const int32_t m0 = ctx->mat_h[lo];
                for (int64_t g = lo; g < hi && uni; g++) {
                    uni = ctx->mat_h[g] == m0 && m0 >= 0;
                }
            }
            (uni ? la : lb).push_back((uint32_t)t);
            if (uni && ctx->mat_has_qm[(size_t)ctx->mat_h[lo]]) {
                ctx->n_tiles_mfma++;
            }
        }
        ctx->n_tiles_mixed = lb.size();
        lb.insert(lb.end(), la.begin(), la.end());
        std::vector<int32_t> lm(lb.size(), -1);
        for (size_t i = ctx->n_tiles_mixed; i < lb.size(); i++) {
            const int64_t lo = (int64_t)(lb[i] * ctx->owned) - ctx->halo;
            const int32_t m0 = ctx->mat_h[lo];
            lm[i] = m0 | (ctx->mat_has_qm[(size_t)m0] ? TILE_MAT_QM : 0);
        }
        if ((rc = upload(ctx, &ctx->tiles_a, lb.data(), lb.size()))) {
            return rc;
        }
        if ((rc = upload(ctx, &ctx->tiles_mat, lm.data(), lm.size()))) {
            return rc;
        }
    }
    if ((rc = upload(ctx, &ctx->tile_src, ts.data(), ts.size()))) {
        return rc;
    }
    if ((rc = upload(ctx, &ctx->tile_rec_ptr, ptr.data(), ptr.size()))) {
        return rc;
    }
    return upload(ctx, &ctx->tile_rec_idx, idx.data(), idx.size());
}

KArgs make_args(const mbh_ctx* ctx, int in_buf, uint64_t n_off,
                uint32_t nsteps);
hipError_t launch_block(const mbh_ctx* ctx, KArgs k);

constexpr uint32_t FP32_MIN_BLOCK = 8;
constexpr double FP32_TIE = 0.015;

/*
 * Autotuner: times every admissible (cells per lane, temporal block) pair
 * on the context's own state and picks the fastest per cell update.  The
 * trial launches read the current buffer, write the other one (discarded)
 * and record nothing, so the simulation state is untouched.
 */
int autotune(mbh_ctx* ctx, uint32_t req_tb, uint32_t req_tile)
{
    uint32_t kform = kernel_form(ctx->algo, ctx->levels, ctx->form);
    if (ctx->algo == MBH_ALGO_NOOP && kform == MBH_FORM_MIXED &&
        ctx->hz_zero && !ctx->generic) {
        kform = KFORM_NOOP_HZ0;
    }
    if (ctx->bloch && ctx->bloch_real && ctx->hz_zero && !ctx->generic) {
        kform = ctx->fp32 ? KFORM_BLOCH2R_MIXED : KFORM_BLOCH2R_ZFORM64;
    }
    ctx->kform = kform;
    std::vector<uint32_t> cs;
    /* cells per lane (tiles of several wavefronts count as their cells
     * per lane of the first) */
    for (uint32_t C : { 1u, 2u, 4u, 8u, 16u }) {
        if (find_kernel(ctx->algo, ctx->levels, kform, C, ctx->order) &&
            (!req_tile || req_tile == 64u * C)) {
            cs.push_back(C);
        }
    }
    if (cs.empty()) {
        return fail(ctx, MBH_ERR_UNSUPPORTED,
                    "no kernel with %u cells per wavefront", req_tile);
    }
    /* FP32 storage rounds the state at every launch boundary, and the
     * error of a run grows with the number of those roundings: automatic
     * blocks of these forms are at least FP32_MIN_BLOCK steps deep (where
     * the run is that long) */
    const uint32_t min_tb = (ctx->fp32 && ctx->n_t >= 4u * FP32_MIN_BLOCK)
        ? FP32_MIN_BLOCK
        : 1u;
    std::vector<std::pair<uint32_t, uint32_t>> cand;
    for (uint32_t C : cs) {
        for (uint32_t tb : { 1u, 2u, 4u, 6u, 8u, 12u, 16u, 20u, 24u, 28u,
                             32u, 48u, 64u }) {
            if (req_tb && tb != req_tb) {
                continue;
            }
            if (tb > MAX_BLOCK || 64u * C < 2u * tb + 8u) {
                continue;
            }
            if (!req_tb && (tb > std::max<uint64_t>(1, ctx->n_t) ||
                            tb < min_tb)) {
                continue;
            }
            cand.emplace_back(C, tb);
        }
    }
    if (cand.empty()) {
        /* requested block too deep for any tile: widest tile, clamped to
         * what it and the kernels (one bit per step of a block in the
         * record mask) allow */
        const uint32_t C = cs.back();
        cand.emplace_back(C, std::min<uint32_t>({req_tb, 32u * C - 4u,
                                                 (uint32_t)MAX_BLOCK}));
    }
    /* cheapest candidates (shallow blocks) first would bias the budget;
     * the list is tried in order until ~0.5 s of trial launches are spent
     * (a single candidate is timed too: its launch time sets the automatic
     * graph size) */
    double best = 1e300, spent_ms = 0.0;
    std::pair<uint32_t, uint32_t> pick = cand[0];
    std::vector<std::pair<double, std::pair<uint32_t, uint32_t>>> timed;
    size_t tried = 0;
    hipLaunchKernelGGL(set_u64, dim3(1), dim3(1), 0, ctx->stream,
                       ctx->n_base, (uint64_t)0);
    for (auto& cd : cand) {
        if (spent_ms > 500.0) {
            break;
        }
        set_geometry(ctx, cd.first, cd.second);
        ctx->launch =
            find_kernel(ctx->algo, ctx->levels, kform, cd.first, ctx->order)
                ->fn;
        int rc = build_tile_tables(ctx, false);
        if (rc) {
            return rc;
        }
        const uint32_t steps = (uint32_t)std::min<uint64_t>(cd.second,
                                                            ctx->n_t);
        KArgs k = make_args(ctx, ctx->cur, 0, steps);
        /* one warm-up launch, timed */
        HIP_TRY(ctx, hipEventRecord(ctx->ev0, ctx->stream));
        HIP_TRY(ctx, launch_block(ctx, k));
        HIP_TRY(ctx, hipEventRecord(ctx->ev1, ctx->stream));
        HIP_TRY(ctx, hipEventSynchronize(ctx->ev1));
        float ms0 = 0.0f;
        HIP_TRY(ctx, hipEventElapsedTime(&ms0, ctx->ev0, ctx->ev1));
        /* about 3 ms of launches per measurement, the faster of two */
        const int reps = std::max(
            1, std::min(64, (int)std::ceil(3.0 / std::max(ms0, 1e-3f))));
        float ms = 1e30f;
        for (int round = 0; round < 2; round++) {
            HIP_TRY(ctx, hipEventRecord(ctx->ev0, ctx->stream));
            for (int i = 0; i < reps; i++) {
                HIP_TRY(ctx, launch_block(ctx, k));
            }
            HIP_TRY(ctx, hipEventRecord(ctx->ev1, ctx->stream));
            HIP_TRY(ctx, hipEventSynchronize(ctx->ev1));
            float mr = 0.0f;
            HIP_TRY(ctx, hipEventElapsedTime(&mr, ctx->ev0, ctx->ev1));
            spent_ms += mr;
            ms = std::min(ms, mr);
        }
        spent_ms += ms0;
        tried++;
        const double per_step = (double)ms / (reps * (double)steps);
        timed.emplace_back(per_step, cd);
        if (per_step < best) {
            best = per_step;
            pick = cd;
        }
    }
    /* FP32 storage: among candidates of the fastest one's tile width
     * within FP32_TIE of it (the spread of the trial timings), the deepest
     * block, whose runs round their state the fewest times.  (A tile of
     * four wavefronts inside that spread has run its full pass 5 % slower
     * than the fastest width.) */
    if (ctx->fp32) {
        const double limit = best * (1.0 + FP32_TIE);
        const uint32_t width = pick.first;
        for (const auto& tc : timed) {
            if (tc.second.first == width && tc.first <= limit &&
                tc.second.second > pick.second) {
                pick = tc.second;
                best = tc.first;
            }
        }
    }
    ctx->tuned_candidates = tried;
    ctx->launch_ms_est = best * pick.second;
    set_geometry(ctx, pick.first, pick.second);
    ctx->launch =
        find_kernel(ctx->algo, ctx->levels, kform, ctx->C, ctx->order)->fn;
    return MBH_OK;
}

KArgs make_args(const mbh_ctx* ctx, int in_buf, uint64_t n_off,
                uint32_t nsteps)
{
    KArgs k{};
    const int ob = 1 - in_buf;
    k.e_in = ctx->state[in_buf][0];
    k.h_in = ctx->state[in_buf][1];
    k.p_in = ctx->state[in_buf][2];
    k.d_in = ctx->state[in_buf][3];
    k.e_out = ctx->state[ob][0];
    k.h_out = ctx->state[ob][1];
    k.p_out = ctx->state[ob][2];
    k.d_out = ctx->state[ob][3];
    k.G = ctx->G;
    k.Gown = owned_cells(ctx);
    k.dstride = ctx->dstride;
    k.mat = ctx->mat;
    k.hzero = ctx->hzero;
    k.fdtd = ctx->fdtd;
    k.qm = ctx->qm;
    k.qstride = ctx->qstride;
    k.yshift = yshift(ctx);
    k.src = ctx->src;
    k.tile_src = ctx->tile_src;
    k.rec = ctx->rec;
    k.tile_rec_ptr = ctx->tile_rec_ptr;
    k.tile_rec_idx = ctx->tile_rec_idx;
    k.n_base = ctx->n_base;
    k.n_off = n_off;
    k.nsteps = nsteps;
    k.halo = ctx->halo;
    k.owned = ctx->owned;
    k.frag = ctx->frag;
    k.frag_stride = ctx->frag_stride;
    return k;
}

/* One temporal block of all tiles (MFMA form: in the order of the tile
 * list, tiles with more than one material first). */
hipError_t launch_block(const mbh_ctx* ctx, KArgs k)
{
    k.tile_list = ctx->split ? ctx->tiles_a : nullptr;
    k.tile_mat = ctx->split ? ctx->tiles_mat : nullptr;
    k.n_mixed = ctx->split ? (uint32_t)ctx->n_tiles_mixed : 0u;
    return ctx->launch(k, ctx->num_tiles, ctx->stream);
}

/*
 * Launches per captured graph for graph_launches == 0.  A graph pays off
 * only where the gaps between launches are a visible share of the time:
 * launches shorter than AUTO_GRAPH_MAX_LAUNCH_MS are grouped into graphs of
 * an even count covering about AUTO_GRAPH_SPAN_MS, provided a full run
 * replays such a graph at least twice; longer launches go one by one.
 */
constexpr double AUTO_GRAPH_MAX_LAUNCH_MS = 0.25;
constexpr double AUTO_GRAPH_SPAN_MS = 2.0;
constexpr uint32_t AUTO_GRAPH_MAX = 128;

uint32_t auto_graph_launches(const mbh_ctx* ctx)
{
    const double t = ctx->launch_ms_est;
    if (!(t > 0.0) || t > AUTO_GRAPH_MAX_LAUNCH_MS) {
        return 1u;
    }
    uint32_t L = 2u * (uint32_t)std::ceil(AUTO_GRAPH_SPAN_MS / (2.0 * t));
    L = std::min(std::max(L, 2u), AUTO_GRAPH_MAX);
    while (L >= 2u && 2ull * L * ctx->tb > ctx->n_t) {
        L -= 2u;
    }
    return L >= 2u ? L : 1u;
}

int build_graph(mbh_ctx* ctx, int start_buf)
{
    hipGraph_t g = nullptr;
    HIP_TRY(ctx, hipStreamBeginCapture(ctx->stream,
                                       hipStreamCaptureModeThreadLocal));
    int buf = start_buf;
    hipError_t le = hipSuccess;
    for (uint32_t j = 0; j < ctx->graph_launches && le == hipSuccess; j++) {
        KArgs k = make_args(ctx, buf, (uint64_t)j * ctx->tb, ctx->tb);
        le = launch_block(ctx, k);
        buf = 1 - buf;
    }
    if (le == hipSuccess) {
        hipLaunchKernelGGL(add_u64, dim3(1), dim3(1), 0, ctx->stream,
                           ctx->n_base,
                           (uint64_t)ctx->graph_launches * ctx->tb);
        le = hipGetLastError();
    }
    hipError_t ce = hipStreamEndCapture(ctx->stream, &g);
    if (le != hipSuccess || ce != hipSuccess) {
        if (g) {
            (void)hipGraphDestroy(g);
        }
        return fail(ctx, MBH_ERR_HIP, "graph capture failed: %s",
                    hipGetErrorString(le != hipSuccess ? le : ce));
    }
    hipError_t ie = hipGraphInstantiate(&ctx->graph[start_buf], g, nullptr,
                                        nullptr, 0);
    (void)hipGraphDestroy(g);
    if (ie != hipSuccess) {
        ctx->graph[start_buf] = nullptr;
        return fail(ctx, MBH_ERR_HIP, "graph instantiation failed: %s",
                    hipGetErrorString(ie));
    }
    return MBH_OK;
}

} /* namespace */

extern "C" {

const char* mbh_version(void) { return "mbsolve-hip 1 gfx950"; }

int mbh_device_count(int* count)
{
    if (!count) {
        return fail(nullptr, MBH_ERR_INVALID, "count is NULL");
    }
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess) {
        *count = 0;
        return MBH_OK;
    }
    *count = n;
    return MBH_OK;
}

void mbh_default_options(mbh_options* opt)
{
    if (!opt) {
        return;
    }
    std::memset(opt, 0, sizeof(*opt));
    opt->form = MBH_FORM_FP64;
    opt->device = -1;
}

int mbh_create(const mbh_problem* runs, uint32_t num_runs,
               const mbh_options* opt_in, mbh_ctx** out)
{
    if (!out) {
        return fail(nullptr, MBH_ERR_INVALID, "out is NULL");
    }
    *out = nullptr;
    if (!runs || num_runs == 0) {
        return fail(nullptr, MBH_ERR_INVALID, "no runs given");
    }
    mbh_options opt;
    mbh_default_options(&opt);
    if (opt_in) {
        opt = *opt_in;
    }
    if (opt.form > MBH_FORM_MFMA64) {
        return fail(nullptr, MBH_ERR_INVALID, "unknown form %u", opt.form);
    }
    if (opt.step_order > MBH_ORDER_RED) {
        return fail(nullptr, MBH_ERR_INVALID, "unknown step order %u",
                    opt.step_order);
    }
    if (opt.step_order == MBH_ORDER_RED && opt.form != MBH_FORM_FP64) {
        return fail(nullptr, MBH_ERR_UNSUPPORTED,
                    "the red step order is implemented for the FP64 form "
                    "only (form %u)", opt.form);
    }
    for (uint32_t r = 0; r < num_runs; r++) {
        int rc = validate(&runs[r], r, &runs[0], opt.form);
        if (rc) {
            return rc;
        }
    }
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) {
        return fail(nullptr, MBH_ERR_NODEVICE, "no HIP device available");
    }

    std::unique_ptr<mbh_ctx> ctx(new mbh_ctx());
    mbh_ctx* c = ctx.get();
    if (opt.device >= 0) {
        if (opt.device >= ndev) {
            return fail(nullptr, MBH_ERR_INVALID, "device %d of %d",
                        opt.device, ndev);
        }
        c->device = opt.device;
        HIP_TRY(nullptr, hipSetDevice(c->device));
    } else {
        HIP_TRY(nullptr, hipGetDevice(&c->device));
    }

    const mbh_problem& p0 = runs[0];
    c->algo = p0.algo;
    c->levels = (p0.algo == MBH_ALGO_NOOP) ? 0 : p0.num_levels;
    c->form = opt.form;
    c->order = opt.step_order;
    c->n_t = p0.num_timesteps;
    c->dref = ref_density_stride(c->algo, c->levels);
    c->DC = c->dref;
    c->zform = (c->algo == MBH_ALGO_CVR_RODR) &&
        (opt.form == MBH_FORM_MIXED || opt.form == MBH_FORM_ZFORM64 ||
         opt.form == MBH_FORM_MFMA || opt.form == MBH_FORM_MFMA64);
    c->fp32 = (opt.form == MBH_FORM_MIXED || opt.form == MBH_FORM_MFMA);
    c->bloch = (c->algo == MBH_ALGO_REG_CAYLEY) && c->levels == 2 &&
        (opt.form == MBH_FORM_MIXED || opt.form == MBH_FORM_ZFORM64);
    if (c->bloch) {
        c->DC = 4;
    }
    c->generic = (opt.flags & MBH_OPT_GENERIC) != 0;
    const size_t esz = c->fp32 ? sizeof(float) : sizeof(double);

    /* super-grid layout */
    uint64_t G = 0, scratch_total = 0;
    uint32_t nmat_total = 0;
    for (uint32_t r = 0; r < num_runs; r++) {
        run_layout L;
        L.base = G;
        L.cells = runs[r].num_gridpoints;
        L.scratch = runs[r].scratch_size;
        L.scratch_off = scratch_total;
        L.dstride_ref = c->dref;
        c->runs.push_back(L);
        G += L.cells + 1;
        scratch_total += runs[r].scratch_size;
        nmat_total += runs[r].num_materials;
    }
    c->G = G;
    c->dstride = (G + 63) & ~uint64_t(63);

    /* per-cell material index and H boundary flags */
    std::vector<int32_t> mat(G, -1);
    std::vector<uint8_t> hz(G, 1);
    std::vector<double> fdtd((size_t)nmat_total * FDTD_STRIDE);
    /* the device copy of a z-form blob gets y* appended (OFF_YEQ) */
    const uint64_t src_stride =
        c->zform ? runs[0].zform_stride : runs[0].qm_stride;
    c->qstride = c->bloch ? BLOCH2_STRIDE
                          : (c->zform ? src_stride + c->dref : src_stride);
    std::vector<double> qm((size_t)nmat_total * c->qstride);
    uint32_t moff = 0;
    for (uint32_t r = 0; r < num_runs; r++) {
        const mbh_problem& p = runs[r];
        const run_layout& L = c->runs[r];
        for (uint64_t i = 0; i < L.cells; i++) {
            mat[L.base + i] = (int32_t)(moff + p.mat_idx[i]);
            hz[L.base + i] = (i == 0) ? 1 : 0;
        }
        for (uint32_t m = 0; m < p.num_materials; m++) {
            for (int j = 0; j < 4; j++) {
                fdtd[(size_t)(moff + m) * FDTD_STRIDE + j] =
                    p.fdtd_consts[(size_t)m * 4 + j];
            }
            fdtd[(size_t)(moff + m) * FDTD_STRIDE + 4] = p.dx_inv;
            const double* q = c->zform ? p.zform_consts : p.qm_consts;
            if (c->bloch) {
                bloch2_constants(q + (size_t)m * p.qm_stride,
                                 &qm[(size_t)(moff + m) * c->qstride]);
            } else if (c->qstride) {
                double* dst = &qm[(size_t)(moff + m) * c->qstride];
                std::memcpy(dst, q + (size_t)m * src_stride,
                            src_stride * sizeof(double));
                if (c->zform) {
                    zform_yeq(c->levels, dst, dst + src_stride);
                }
            }
        }
        moff += p.num_materials;
    }

    if (c->bloch) {
        /* algo_bloch2r: beta_y, d alpha0 / d e and the p coefficients of
         * tau and m_z exactly zero in every material with a density, and
         * every material lossless (fac_a exactly 1) */
        c->bloch_real = true;
        for (size_t mi = 0; mi + BLOCH2_STRIDE <= qm.size();
             mi += BLOCH2_STRIDE) {
            const double* o = &qm[mi];
            if (o[0] != 0.0 && (o[2] != 0.0 || o[5] != 0.0 || o[8] != 0.0 ||
                                o[14] != 0.0 || o[17] != 0.0)) {
                c->bloch_real = false;
            }
        }
        for (size_t mi = 0; mi < fdtd.size(); mi += FDTD_STRIDE) {
            if (fdtd[mi] != 1.0) {
                c->bloch_real = false;
            }
        }
    }

    /* kernel and geometry */
    /* graphs hold an even number of launches so that the ping-pong buffer
     * parity is the same before and after a replay */
    c->graph_launches =
        (opt.graph_launches > 1) ? ((opt.graph_launches + 1u) & ~1u) : 1u;
    const bool graph_auto = (opt.graph_launches == 0);
    if (!find_kernel(c->algo, c->levels,
                     kernel_form(c->algo, c->levels, c->form), 1, c->order)) {
        return fail(nullptr, MBH_ERR_UNSUPPORTED,
                    "no kernel for algo %u, %u levels, form %u", c->algo,
                    c->levels, c->form);
    }
    if (c->zform) {
        for (uint32_t i = 0; i < z_table_n; i++) {
            if (z_table[i].levels == c->levels &&
                z_table[i].fp32 == (c->fp32 ? 1u : 0u)) {
                c->to_y = z_table[i].to_y;
                c->from_y = z_table[i].from_y;
            }
        }
        if (!c->to_y) {
            return fail(nullptr, MBH_ERR_UNSUPPORTED,
                        "no z-form conversion for %u levels", c->levels);
        }
    }

    HIP_TRY(c, hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
    HIP_TRY(c, hipEventCreate(&c->ev0));
    HIP_TRY(c, hipEventCreate(&c->ev1));

    int rc;
#define RC(x)                                                                 \
    do {                                                                      \
        if ((rc = (x)) != MBH_OK) {                                           \
            g_last_error = c->err;                                            \
            return rc;                                                        \
        }                                                                     \
    } while (0)

    RC(upload(c, &c->mat, mat.data(), mat.size()));
    RC(upload(c, &c->hzero, hz.data(), hz.size()));
    RC(upload(c, &c->fdtd, fdtd.data(), fdtd.size()));
    RC(upload(c, &c->qm, qm.data(), qm.size()));
    if (c->form == MBH_FORM_MFMA || c->form == MBH_FORM_MFMA64) {
        const uint64_t nm = qm.size() / std::max<uint64_t>(1, c->qstride);
        c->frag_stride = zm_bytes(c->levels);
        std::vector<unsigned char> fr(nm * c->frag_stride);
        for (uint64_t mi = 0; mi < nm; mi++) {
            zm_build_levels(c->levels, &qm[mi * c->qstride],
                            &fr[mi * c->frag_stride]);
        }
        unsigned char* dfr = nullptr;
        RC(upload(c, &dfr, fr.data(), fr.size()));
        c->frag = dfr;
    }

    /* state buffers: e, h, p [G], density [DC][dstride] */
    for (int b = 0; b < 2; b++) {
        RC(dalloc(c, &c->state[b][0], G * esz));
        RC(dalloc(c, &c->state[b][1], G * esz));
        RC(dalloc(c, &c->state[b][2], G * esz));
        RC(dalloc(c, &c->state[b][3],
                  std::max<uint64_t>(1, c->DC) * c->dstride * esz));
        for (int j = 0; j < 4; j++) {
            HIP_TRY(c, hipMemset(c->state[b][j], 0,
                                 (j < 3 ? G : std::max<uint64_t>(1, c->DC) *
                                          c->dstride) *
                                     esz));
        }
    }
    c->state_bytes = (3 * G + (uint64_t)c->DC * c->dstride) * esz;

    /* initial state */
    {
        std::vector<double> e(G, 0.0), h(G, 0.0);
        std::vector<double> d((size_t)c->dref * c->dstride, 0.0);
        for (uint32_t r = 0; r < num_runs; r++) {
            const mbh_problem& p = runs[r];
            const run_layout& L = c->runs[r];
            for (uint64_t i = 0; i < L.cells; i++) {
                e[L.base + i] = p.e_init[i];
                h[L.base + i] = p.h_init[i];
                for (uint32_t j = 0; j < c->dref; j++) {
                    d[(size_t)j * c->dstride + L.base + i] =
                        p.d_init[i * c->dref + j];
                }
            }
            h[L.base + L.cells] = p.h_init[L.cells];
        }
        c->hz_zero = true;
        for (uint64_t g = 0; g < G; g++) {
            if (hz[g] && h[g] != 0.0) {
                c->hz_zero = false;
                break;
            }
        }
        if (c->bloch) {
            /* rho00, rho11 and rho01 of the Hermitian state, converted to
             * the carried Bloch state of the cell's material */
            std::vector<double> sel((size_t)c->DC * c->dstride, 0.0);
            for (uint64_t g = 0; g < G; g++) {
                if (mat[g] < 0) {
                    continue;
                }
                double rho[4], sig[4];
                for (uint32_t j = 0; j < 4; j++) {
                    rho[j] = d[(size_t)BLOCH_FROM_REF[j] * c->dstride + g];
                }
                bloch2_to_sigma(&qm[(size_t)mat[g] * c->qstride], rho, sig);
                for (uint32_t j = 0; j < 4; j++) {
                    sel[(size_t)j * c->dstride + g] = sig[j];
                }
            }
            d.swap(sel);
            c->mat_h = mat;
            c->qm_h = qm;
        }
        if (c->fp32) {
            std::vector<float> ef(e.begin(), e.end()), hf(h.begin(), h.end());
            HIP_TRY(c, hipMemcpy(c->state[0][0], ef.data(), G * esz,
                                 hipMemcpyHostToDevice));
            HIP_TRY(c, hipMemcpy(c->state[0][1], hf.data(), G * esz,
                                 hipMemcpyHostToDevice));
        } else {
            HIP_TRY(c, hipMemcpy(c->state[0][0], e.data(), G * esz,
                                 hipMemcpyHostToDevice));
            HIP_TRY(c, hipMemcpy(c->state[0][1], h.data(), G * esz,
                                 hipMemcpyHostToDevice));
        }
        if (c->dref) {
            if (c->zform) {
                double* dtmp = nullptr;
                RC(upload(c, &dtmp, d.data(), d.size()));
                ZArgs za{ dtmp,    c->state[0][3], G, c->dstride, c->mat,
                          c->qm,   c->qstride,     yshift(c), 0, G,
                          nullptr };
                HIP_TRY(c, c->to_y(za, c->stream));
                HIP_TRY(c, hipStreamSynchronize(c->stream));
            } else if (c->fp32) {
                std::vector<float> df(d.begin(), d.end());
                HIP_TRY(c, hipMemcpy(c->state[0][3], df.data(),
                                     df.size() * sizeof(float),
                                     hipMemcpyHostToDevice));
            } else {
                HIP_TRY(c, hipMemcpy(c->state[0][3], d.data(),
                                     d.size() * sizeof(double),
                                     hipMemcpyHostToDevice));
            }
        }
        c->cur = 0;
    }

    /* a copy of the initial state for mbh_reset */
    c->state_elems[0] = c->state_elems[1] = c->state_elems[2] = G;
    c->state_elems[3] = std::max<uint64_t>(1, c->DC) * c->dstride;
    for (int j = 0; j < 4; j++) {
        RC(dalloc(c, &c->init[j], c->state_elems[j] * esz));
        HIP_TRY(c, hipMemcpy(c->init[j], c->state[0][j],
                             c->state_elems[j] * esz,
                             hipMemcpyDeviceToDevice));
    }

    /* source tables, sorted by cell; stable, so sources at one cell keep
     * the reference's list order */
    {
        uint64_t total = 0;
        for (uint32_t r = 0; r < num_runs; r++) {
            total += (uint64_t)runs[r].num_sources * c->n_t;
        }
        RC(dalloc(c, (void**)&c->src_data, total * sizeof(double)));
        uint64_t off = 0;
        for (uint32_t r = 0; r < num_runs; r++) {
            const mbh_problem& p = runs[r];
            for (uint32_t k = 0; k < p.num_sources; k++) {
                HIP_TRY(c, hipMemcpy(c->src_data + off,
                                     p.source_data + (size_t)k * c->n_t,
                                     c->n_t * sizeof(double),
                                     hipMemcpyHostToDevice));
                dev_source s{};
                s.gx = c->runs[r].base + p.sources[k].x_idx;
                s.data = c->src_data + off;
                s.type = p.sources[k].type;
                c->srcs.push_back(s);
                off += c->n_t;
            }
        }
        std::stable_sort(c->srcs.begin(), c->srcs.end(),
                         [](const dev_source& a, const dev_source& b) {
                             return a.gx < b.gx;
                         });
        RC(upload(c, &c->src, c->srcs.data(), c->srcs.size()));
    }

    /* record scratch and descriptors */
    {
        c->rec_total = std::max<uint64_t>(1, scratch_total);
        RC(dalloc(c, (void**)&c->rec_buf,
                  std::max<uint64_t>(1, scratch_total) * sizeof(double)));
        HIP_TRY(c, hipMemset(c->rec_buf, 0,
                             std::max<uint64_t>(1, scratch_total) *
                                 sizeof(double)));
        for (uint32_t r = 0; r < num_runs; r++) {
            const mbh_problem& p = runs[r];
            const run_layout& L = c->runs[r];
            for (uint32_t k = 0; k < p.num_records; k++) {
                const mbh_record& R = p.records[k];
                dev_record d{};
                d.gbegin = L.base + R.position_idx;
                d.cols = R.cols;
                d.rows = R.rows;
                d.ratio = R.interval_ratio;
                d.dst = c->rec_buf + L.scratch_off + R.offset_real;
                d.type = R.type;
                d.row_idx = R.row_idx;
                d.col_idx = R.col_idx;
                c->recs.push_back(d);
            }
        }
        RC(upload(c, &c->rec, c->recs.data(), c->recs.size()));
    }

    RC(dalloc(c, (void**)&c->n_base, sizeof(uint64_t)));
    HIP_TRY(c, hipMemset(c->n_base, 0, sizeof(uint64_t)));
    HIP_TRY(c, hipDeviceSynchronize());

    /* MFMA form: the tile list puts tiles with more than one material
     * (VALU path inside the kernel) first */
    if (c->form == MBH_FORM_MFMA || c->form == MBH_FORM_MFMA64) {
        c->split = true;
        c->mat_h = mat;
        c->mat_has_qm.resize(qm.size() / std::max<uint64_t>(1, c->qstride));
        for (size_t mi = 0; mi < c->mat_has_qm.size(); mi++) {
            c->mat_has_qm[mi] = qm[mi * c->qstride] != 0.0 ? 1 : 0;
        }
    }

    /* launch geometry: measured, then the final per-tile tables */
    RC(autotune(c, opt.temporal_block, opt.tile_cells));
    RC(build_tile_tables(c, true));
    if (graph_auto) {
        c->graph_launches = auto_graph_launches(c);
    }
    HIP_TRY(c, hipDeviceSynchronize());
#undef RC
    *out = ctx.release();
    return MBH_OK;
}

int mbh_run(mbh_ctx* ctx, uint64_t n_begin, uint64_t n_end)
{
    if (!ctx) {
        return fail(nullptr, MBH_ERR_INVALID, "ctx is NULL");
    }
    if (n_begin > n_end || n_end > ctx->n_t) {
        return fail(ctx, MBH_ERR_INVALID,
                    "step range [%llu, %llu) outside [0, %llu)",
                    (unsigned long long)n_begin, (unsigned long long)n_end,
                    (unsigned long long)ctx->n_t);
    }
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    HIP_TRY(ctx, hipEventRecord(ctx->ev0, ctx->stream));
    hipLaunchKernelGGL(set_u64, dim3(1), dim3(1), 0, ctx->stream,
                       ctx->n_base, n_begin);
    HIP_TRY(ctx, hipGetLastError());

    uint64_t base = n_begin; /* value *n_base will hold at this point */
    uint64_t n = n_begin;
    uint64_t launches = 0;
    const uint64_t per_graph = (uint64_t)ctx->graph_launches * ctx->tb;
    while (n < n_end) {
        const uint64_t left = n_end - n;
        if (ctx->graph_launches > 1 && left >= per_graph) {
            if (!ctx->graph[ctx->cur]) {
                int rc = build_graph(ctx, ctx->cur);
                if (rc) {
                    return rc;
                }
            }
            HIP_TRY(ctx, hipGraphLaunch(ctx->graph[ctx->cur], ctx->stream));
            /* graph_launches is even: the buffer parity is unchanged */
            n += per_graph;
            base += per_graph;
            launches += ctx->graph_launches;
            continue;
        }
        const uint32_t steps = (uint32_t)std::min<uint64_t>(ctx->tb, left);
        KArgs k = make_args(ctx, ctx->cur, n - base, steps);
        HIP_TRY(ctx, launch_block(ctx, k));
        ctx->cur = 1 - ctx->cur;
        n += steps;
        launches++;
    }
    if (n_begin == 0 && n_end > 0) {
        /* step 0 has read h[N] of the last run (the trailing pseudo cell,
         * never written back by a tile); the boundary condition holds it
         * at zero from then on */
        const uint64_t g = ctx->G - 1;
        if (ctx->fp32) {
            hipLaunchKernelGGL(zero_last_h<float>, dim3(1), dim3(1), 0,
                               ctx->stream, (float*)ctx->state[0][1],
                               (float*)ctx->state[1][1], g);
        } else {
            hipLaunchKernelGGL(zero_last_h<double>, dim3(1), dim3(1), 0,
                               ctx->stream, (double*)ctx->state[0][1],
                               (double*)ctx->state[1][1], g);
        }
        HIP_TRY(ctx, hipGetLastError());
    }
    HIP_TRY(ctx, hipEventRecord(ctx->ev1, ctx->stream));
    ctx->timed = true;
    ctx->last_launches = launches;
    return MBH_OK;
}

int mbh_reset(mbh_ctx* ctx)
{
    if (!ctx) {
        return fail(nullptr, MBH_ERR_INVALID, "ctx is NULL");
    }
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    const size_t esz = ctx->fp32 ? sizeof(float) : sizeof(double);
    for (int j = 0; j < 4; j++) {
        HIP_TRY(ctx, hipMemcpyAsync(ctx->state[0][j], ctx->init[j],
                                    ctx->state_elems[j] * esz,
                                    hipMemcpyDeviceToDevice, ctx->stream));
    }
    hipLaunchKernelGGL(fill_zero_f64, dim3(256), dim3(256), 0, ctx->stream,
                       ctx->rec_buf, ctx->rec_total);
    HIP_TRY(ctx, hipGetLastError());
    ctx->cur = 0;
    return MBH_OK;
}

int mbh_synchronize(mbh_ctx* ctx)
{
    if (!ctx) {
        return fail(nullptr, MBH_ERR_INVALID, "ctx is NULL");
    }
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    return MBH_OK;
}

int mbh_elapsed_ms(mbh_ctx* ctx, double* ms)
{
    if (!ctx || !ms) {
        return fail(ctx, MBH_ERR_INVALID, "ctx or ms is NULL");
    }
    if (!ctx->timed) {
        return fail(ctx, MBH_ERR_INVALID, "no mbh_run recorded yet");
    }
    HIP_TRY(ctx, hipEventSynchronize(ctx->ev1));
    float f = 0.0f;
    HIP_TRY(ctx, hipEventElapsedTime(&f, ctx->ev0, ctx->ev1));
    *ms = (double)f;
    return MBH_OK;
}

int mbh_fetch_records(mbh_ctx* ctx, uint32_t run, double* dst,
                      uint64_t count)
{
    if (!ctx || !dst) {
        return fail(ctx, MBH_ERR_INVALID, "ctx or dst is NULL");
    }
    if (run >= ctx->runs.size() || count > ctx->runs[run].scratch) {
        return fail(ctx, MBH_ERR_INVALID,
                    "run %u / count %llu out of range", run,
                    (unsigned long long)count);
    }
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    HIP_TRY(ctx, hipMemcpy(dst, ctx->rec_buf + ctx->runs[run].scratch_off,
                           count * sizeof(double), hipMemcpyDeviceToHost));
    return MBH_OK;
}

int mbh_fetch_state(mbh_ctx* ctx, uint32_t run, double* e, double* h,
                    double* p, double* d)
{
    if (!ctx) {
        return fail(nullptr, MBH_ERR_INVALID, "ctx is NULL");
    }
    if (run >= ctx->runs.size()) {
        return fail(ctx, MBH_ERR_INVALID, "run %u out of range", run);
    }
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    const run_layout& L = ctx->runs[run];
    const size_t esz = ctx->fp32 ? sizeof(float) : sizeof(double);
    auto get = [&](const void* src, uint64_t first, uint64_t count,
                   double* out) -> int {
        if (ctx->fp32) {
            std::vector<float> tmp(count);
            HIP_TRY(ctx, hipMemcpy(tmp.data(),
                                   (const char*)src + first * esz,
                                   count * esz, hipMemcpyDeviceToHost));
            for (uint64_t i = 0; i < count; i++) {
                out[i] = tmp[i];
            }
        } else {
            HIP_TRY(ctx, hipMemcpy(out, (const char*)src + first * esz,
                                   count * esz, hipMemcpyDeviceToHost));
        }
        return MBH_OK;
    };
    int rc;
    void* const* S = ctx->state[ctx->cur];
    if (e && (rc = get(S[0], L.base, L.cells, e))) {
        return rc;
    }
    if (h && (rc = get(S[1], L.base, L.cells + 1, h))) {
        return rc;
    }
    if (p && (rc = get(S[2], L.base, L.cells, p))) {
        return rc;
    }
    if (d && ctx->dref) {
        std::vector<double> comp(L.cells);
        if (ctx->bloch) {
            /* Hermitian 2x2 from the carried Bloch state */
            std::vector<double> sig((size_t)4 * L.cells);
            for (uint32_t j = 0; j < 4; j++) {
                if ((rc = get(S[3], (uint64_t)j * ctx->dstride + L.base,
                              L.cells, sig.data() + (size_t)j * L.cells))) {
                    return rc;
                }
            }
            std::fill(d, d + L.cells * ctx->dref, 0.0);
            for (uint64_t i = 0; i < L.cells; i++) {
                const int32_t m = ctx->mat_h[L.base + i];
                double sv[4], rho[4];
                for (uint32_t j = 0; j < 4; j++) {
                    sv[j] = sig[(size_t)j * L.cells + i];
                }
                bloch2_from_sigma(&ctx->qm_h[(size_t)m * ctx->qstride], sv,
                                  rho);
                double* di = d + i * ctx->dref;
                di[0] = rho[0];
                di[6] = rho[1];
                di[2] = rho[2];
                di[3] = rho[3];
                di[4] = rho[2];
                di[5] = -rho[3];
            }
        } else if (ctx->zform) {
            double* dtmp = nullptr;
            rc = dalloc(ctx, (void**)&dtmp,
                        (size_t)ctx->dref * ctx->dstride * sizeof(double));
            if (rc) {
                return rc;
            }
            ZArgs za{ dtmp,         S[3],        ctx->G,
                      ctx->dstride, ctx->mat,    ctx->qm,
                      ctx->qstride, yshift(ctx), 0,
                      ctx->G,       nullptr };
            HIP_TRY(ctx, ctx->from_y(za, ctx->stream));
            HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
            for (uint32_t j = 0; j < ctx->dref; j++) {
                HIP_TRY(ctx, hipMemcpy(comp.data(),
                                       dtmp + (size_t)j * ctx->dstride +
                                           L.base,
                                       L.cells * sizeof(double),
                                       hipMemcpyDeviceToHost));
                for (uint64_t i = 0; i < L.cells; i++) {
                    d[i * ctx->dref + j] = comp[i];
                }
            }
            ctx->allocs.pop_back();
            (void)hipFree(dtmp);
        } else {
            for (uint32_t j = 0; j < ctx->dref; j++) {
                if ((rc = get(S[3], (uint64_t)j * ctx->dstride + L.base,
                              L.cells, comp.data()))) {
                    return rc;
                }
                for (uint64_t i = 0; i < L.cells; i++) {
                    d[i * ctx->dref + j] = comp[i];
                }
            }
        }
    }
    return MBH_OK;
}

int mbh_fetch_state_noalloc(mbh_ctx* ctx, uint32_t run, double* e, double* h,
                            double* p, double* d)
{
    if (!ctx) {
        return fail(nullptr, MBH_ERR_INVALID, "ctx is NULL");
    }
    if (run >= ctx->runs.size()) {
        return fail(ctx, MBH_ERR_INVALID, "run %u out of range", run);
    }
    const run_layout& L = ctx->runs[run];
    const uint64_t cap = MBH_NOALLOC_SCRATCH_BYTES / sizeof(double);
    const uint64_t need =
        std::max<uint64_t>(L.cells + 1,
                           (d && ctx->dref) ? L.cells * ctx->dref : 0);
    if (need > cap) {
        return fail(ctx, MBH_ERR_INVALID,
                    "run %u needs %llu bytes of staging, more than the "
                    "%llu of mbh_fetch_state_noalloc (use mbh_fetch_state)",
                    run, (unsigned long long)(need * sizeof(double)),
                    (unsigned long long)MBH_NOALLOC_SCRATCH_BYTES);
    }
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    double* stage = nullptr;
    int rc = noalloc_scratch(ctx, &stage);
    if (rc) {
        return rc;
    }
    /* everything on the context's stream, after the queued steps */
    const dim3 grid(256), block(256);
    void* const* S = ctx->state[ctx->cur];
    auto copy_out = [&](double* dst, const double* src,
                        uint64_t count) -> int {
        HIP_TRY(ctx, hipMemcpyAsync(dst, src, count * sizeof(double),
                                    hipMemcpyDeviceToHost, ctx->stream));
        HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
        return MBH_OK;
    };
    auto field = [&](const void* src, uint64_t count, double* dst) -> int {
        const uint64_t first = L.base;
        if (ctx->fp32) {
            hipLaunchKernelGGL(widen_range<float>, grid, block, 0,
                               ctx->stream, (const float*)src + first, stage,
                               count);
            HIP_TRY(ctx, hipGetLastError());
            return copy_out(dst, stage, count);
        }
        /* FP64 state: straight from the state buffer */
        return copy_out(dst, (const double*)src + first, count);
    };
    if (e && (rc = field(S[0], L.cells, e))) {
        return rc;
    }
    if (h && (rc = field(S[1], L.cells + 1, h))) {
        return rc;
    }
    if (p && (rc = field(S[2], L.cells, p))) {
        return rc;
    }
    if (d && ctx->dref) {
        if (ctx->bloch) {
            if (ctx->fp32) {
                hipLaunchKernelGGL(bloch_to_ref<float>, grid, block, 0,
                                   ctx->stream, (const float*)S[3],
                                   ctx->dstride, L.base, L.cells, ctx->mat,
                                   ctx->qm, ctx->qstride, stage);
            } else {
                hipLaunchKernelGGL(bloch_to_ref<double>, grid, block, 0,
                                   ctx->stream, (const double*)S[3],
                                   ctx->dstride, L.base, L.cells, ctx->mat,
                                   ctx->qm, ctx->qstride, stage);
            }
        } else if (ctx->zform) {
            ZArgs za{ nullptr,      S[3],        ctx->G,
                      ctx->dstride, ctx->mat,    ctx->qm,
                      ctx->qstride, yshift(ctx), L.base,
                      L.cells,      stage };
            HIP_TRY(ctx, ctx->from_y(za, ctx->stream));
        } else if (ctx->fp32) {
            hipLaunchKernelGGL(gather_cell_major<float>, grid, block, 0,
                               ctx->stream, (const float*)S[3], ctx->dstride,
                               L.base, L.cells, ctx->dref, stage);
        } else {
            hipLaunchKernelGGL(gather_cell_major<double>, grid, block, 0,
                               ctx->stream, (const double*)S[3],
                               ctx->dstride, L.base, L.cells, ctx->dref,
                               stage);
        }
        HIP_TRY(ctx, hipGetLastError());
        if ((rc = copy_out(d, stage, L.cells * ctx->dref))) {
            return rc;
        }
    }
    return MBH_OK;
}

int mbh_get_info(const mbh_ctx* ctx, mbh_info* info)
{
    if (!ctx || !info) {
        return fail(nullptr, MBH_ERR_INVALID, "ctx or info is NULL");
    }
    std::memset(info, 0, sizeof(*info));
    info->form = ctx->form;
    info->temporal_block = ctx->tb;
    info->tile_cells = 64u * ctx->C;
    info->owned_cells = ctx->owned;
    info->num_tiles = ctx->num_tiles;
    uint64_t total = 0;
    for (auto& r : ctx->runs) {
        total += r.cells;
    }
    info->total_cells = total;
    info->launches = ctx->last_launches;
    info->state_bytes = ctx->state_bytes;
    info->graph_launches = ctx->graph_launches;
    info->step_order = ctx->order;
    info->mfma_tiles = (uint32_t)ctx->n_tiles_mfma;
    info->flags = 0u;
    if (ctx->kform == KFORM_NOOP_HZ0) {
        info->flags |= MBH_INFO_HZ_ZERO;
    }
    if (ctx->kform == KFORM_BLOCH2R_MIXED ||
        ctx->kform == KFORM_BLOCH2R_ZFORM64) {
        info->flags |= MBH_INFO_BLOCH_REAL;
    }
    return MBH_OK;
}

void mbh_destroy(mbh_ctx* ctx) { delete ctx; }

const char* mbh_last_error(const mbh_ctx* ctx)
{
    if (ctx) {
        return ctx->err.c_str();
    }
    return g_last_error.c_str();
}

} /* extern "C" */
