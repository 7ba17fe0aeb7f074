/*
 * Host set-up of an FDTD Maxwell-Bloch run (see fdtd_setup.hpp).
 */
#include <cmath>
#include <cstring>
#include <iostream>
#include <limits>
#include <map>
#include <stdexcept>
#include <mbsolve/lib/internal/cv_representation.hpp>
#include <mbsolve/lib/internal/fdtd_setup.hpp>
#include <mbsolve/lib/internal/numerics.hpp>
#include <mbsolve/lib/material.hpp>

namespace mbsolve {

using num::rmat;

void
init_fdtd_simulation(std::shared_ptr<const device> dev,
                     std::shared_ptr<scenario> scen, real courant)
{
    if (scen->get_num_gridpoints() == 0) {
        throw std::invalid_argument("Invalid scenario.");
    }
    const real velocity =
        1.0 / std::sqrt(MU0 * EPS0 * dev->get_minimum_permittivity());
    const unsigned int n_x = scen->get_num_gridpoints();
    unsigned int n_t;
    if (n_x == 1) {
        /* 0D: one infinitely large grid point (approx. one light year) and
         * the user's step count */
        if (dev->get_length() > 0.0) {
            std::cout << "Warning: Device with length = " << dev->get_length()
                      << " simulated using only one grid point!" << std::endl;
        }
        scen->set_gridpoint_size(1e13);
        n_t = scen->get_num_timesteps();
        if (n_t < 2) {
            throw std::invalid_argument("Invalid scenario.");
        }
    } else {
        const real d_x = dev->get_length() / (n_x - 1);
        scen->set_gridpoint_size(d_x);
        const real d_t = courant * d_x / velocity;
        n_t = std::ceil(scen->get_endtime() / d_t) + 1;
        scen->set_num_timesteps(n_t);
    }
    if (scen->get_endtime() > std::numeric_limits<real>::min()) {
        scen->set_timestep_size(scen->get_endtime() / (n_t - 1));
    } else {
        /* zero end time: one step of approx. one century */
        scen->set_num_timesteps(1);
        scen->set_timestep_size(3.154e10);
    }
}

sim_constants_fdtd
get_fdtd_constants(std::shared_ptr<const device> /* dev */,
                   std::shared_ptr<const scenario> scen,
                   std::shared_ptr<const material> mat)
{
    sim_constants_fdtd sc;
    /* losses as conductivity */
    const real sigma = std::sqrt(EPS0 * mat->get_rel_permittivity() /
                                 (MU0 * mat->get_rel_permeability())) *
        mat->get_losses() * 2.0;
    const real dt = scen->get_timestep_size();
    const real epsilon = EPS0 * mat->get_rel_permittivity();
    const real x = (sigma * dt) / (2.0 * epsilon);
    sc.fac_a = (1.0 - x) / (1.0 + x);
    sc.fac_b = dt / epsilon / (1.0 + x);
    sc.fac_c =
        dt / (MU0 * mat->get_rel_permeability() * scen->get_gridpoint_size());
    sc.gamma = mat->get_overlap_factor();
    return sc;
}

namespace {

void check_levels(std::shared_ptr<const qm_description> qm, unsigned int n)
{
    if (qm->get_num_levels() != n) {
        throw std::invalid_argument("Number of energy levels does not "
                                    "match selected solver!");
    }
}

} /* namespace */

std::vector<real>
cayley_constants(std::shared_ptr<const qm_description> qm, real dt,
                 unsigned int N)
{
    std::vector<real> b(MBH_CAYLEY_STRIDE(N), 0.0);
    b[2] = dt;
    if (!qm) {
        return b; /* vacuum: has_qm = 0, all constants zero */
    }
    check_levels(qm, N);
    auto sop =
        std::dynamic_pointer_cast<qm_lindblad_relaxation>(
            qm->get_relaxation_superop());
    if (!sop) {
        throw std::invalid_argument("Unsupported superoperator!");
    }
    b[0] = 1.0;
    b[1] = qm->get_carrier_density();
    /* -i / hbar * dt as the complex scalar (-0, (-1/hbar) dt) */
    b[3] = (-1.0 / HBAR) * dt;
    b[4] = -1.0 / HBAR;
    const size_t oH0 = 5, oMU = oH0 + 2 * N * N, od1 = oMU + 2 * N * N,
                 od2 = od1 + N * N, oD1 = od2 + N * N, oD2 = oD1 + N * N;
    const auto H0 = cv_representation::to_matrix(qm->get_hamiltonian(), N);
    const auto mu =
        cv_representation::to_matrix(qm->get_dipole_operator(), N);
    for (unsigned int i = 0; i < N * N; i++) {
        b[oH0 + 2 * i] = H0.a[i].real();
        b[oH0 + 2 * i + 1] = H0.a[i].imag();
        b[oMU + 2 * i] = mu.a[i].real();
        b[oMU + 2 * i + 1] = mu.a[i].imag();
    }
    /* population part: scattering matrix and exp(d1 dt / 2) */
    rmat d1(N, N);
    const auto& s = sop->get_scattering_rates();
    for (unsigned int r = 0; r < N; r++) {
        for (unsigned int c = 0; c < N; c++) {
            d1(r, c) = s[r][c];
        }
    }
    rmat arg(N, N);
    for (unsigned int i = 0; i < N * N; i++) {
        arg.a[i] = d1.a[i] * dt / 2;
    }
    const rmat D1 = num::expm(arg);
    /* dephasing part: symmetric rate matrix, elementwise exp */
    rmat d2(N, N);
    const auto& deph = sop->get_dephasing_rates();
    unsigned int row = 0, col = 1;
    for (size_t i = 0; i < deph.size() && col < N; i++) {
        d2(row, col) = deph[i];
        d2(col, row) = deph[i];
        if (row == col - 1) {
            row = 0;
            col++;
        } else {
            row++;
        }
    }
    for (unsigned int i = 0; i < N * N; i++) {
        b[od1 + i] = d1.a[i];
        b[od2 + i] = d2.a[i];
        b[oD1 + i] = D1.a[i];
        b[oD2 + i] = std::exp(d2.a[i] * dt / 2);
    }
    return b;
}

void
rodr_constants(std::shared_ptr<const qm_description> qm, real dt,
               unsigned int N, std::vector<real>& blob,
               std::vector<real>* zblob)
{
    const unsigned int K = N * N - 1, B = K / 2;
    blob.assign(MBH_RODR_STRIDE(N), 0.0);
    if (zblob) {
        zblob->assign(MBH_ZFORM_STRIDE(N), 0.0);
    }
    if (!qm) {
        return; /* vacuum: has_qm = 0 */
    }
    check_levels(qm, N);
    cv_representation cvr(qm);

    const real M_CP = 0.5 * qm->get_carrier_density();
    const rmat& G = cvr.get_relaxation_superop();
    const rmat& M0 = cvr.get_hamiltonian();
    const rmat M = M0 + G;
    rmat U(K, K);
    for (unsigned int i = 0; i < K * K; i++) {
        U.a[i] = -cvr.get_dipole_operator().a[i];
    }
    const rmat U2 = U * U;
    real theta_1 = 0.0;
    if (K >= 3) {
        theta_1 = std::sqrt(std::pow(U(0, 1), 2) + std::pow(U(0, 2), 2) +
                            std::pow(U(1, 2), 2));
    }
    rmat arg(K, K);
    for (unsigned int i = 0; i < K * K; i++) {
        arg.a[i] = M.a[i] * dt / 2;
    }
    const rmat A1 = num::expm(arg);

    const num::skew_blocks sb = num::skew_decompose(U);
    const unsigned int nth = std::min<unsigned int>(sb.theta.size(), B);

    const std::vector<real>& v = cvr.get_dipole_operator_vec();
    const std::vector<real>& d_eq = cvr.get_equilibrium_vec();
    std::vector<real> d_in(K, 0.0);
    bool deq_zero = true;
    for (real x : d_eq) {
        if (std::abs(x) > std::numeric_limits<real>::epsilon()) {
            deq_zero = false;
        }
    }
    if (!deq_zero) {
        d_in = num::solve_full_pivot(M, d_eq);
        const std::vector<real> r = num::matvec(M, d_in);
        real en = 0.0, rn = 0.0;
        for (unsigned int i = 0; i < K; i++) {
            en += (r[i] - d_eq[i]) * (r[i] - d_eq[i]);
            rn += d_eq[i] * d_eq[i];
        }
        if (std::sqrt(en) / std::sqrt(rn) > 1e-3) {
            throw std::invalid_argument("Time-indepent matrix not "
                                        "invertible!");
        }
    }

    /* reference-form blob */
    blob[0] = 1.0;
    blob[1] = M_CP;
    blob[2] = dt;
    blob[3] = theta_1;
    blob[4] = (real)nth;
    size_t o = 5;
    for (unsigned int i = 0; i < K * K; i++) {
        blob[o + i] = M.a[i];
        blob[o + K * K + i] = U.a[i];
        blob[o + 2 * K * K + i] = U2.a[i];
        blob[o + 3 * K * K + i] = A1.a[i];
    }
    o += 4 * K * K;
    for (unsigned int i = 0; i < K; i++) {
        blob[o + i] = v[i];
        blob[o + K + i] = d_eq[i];
        blob[o + 2 * K + i] = d_in[i];
    }
    o += 3 * K;
    const size_t oth = o, oC1 = oth + B, oC2 = oC1 + (size_t)B * K * K;
    for (unsigned int j = 0; j < nth; j++) {
        blob[oth + j] = sb.theta[j];
        /* C1 = Q B_j Q^T = q2 q1^T - q1 q2^T, C2 = Q B_j^2 Q^T =
         * -(q1 q1^T + q2 q2^T) */
        for (unsigned int a = 0; a < K; a++) {
            const real q1a = sb.Q(a, 2 * j), q2a = sb.Q(a, 2 * j + 1);
            for (unsigned int c = 0; c < K; c++) {
                const real q1c = sb.Q(c, 2 * j), q2c = sb.Q(c, 2 * j + 1);
                blob[oC1 + (size_t)j * K * K + a * K + c] =
                    q2a * q1c - q1a * q2c;
                blob[oC2 + (size_t)j * K * K + a * K + c] =
                    -(q1a * q1c + q2a * q2c);
            }
        }
    }

    if (!zblob) {
        return;
    }
    /* rotated basis: y = Q^T d1,  W = Q^T A1^2 Q,  c = Q^T (A1^2 - I) d_in,
     * p = a.r + b with a^T = M_CP v^T M A1 Q,
     * b = M_CP v^T (M (A1 d_in - d_in) + d_eq),
     * d_after = P y' + o with P = A1 Q W^-1, o = (A1 - I) d_in - P c */
    const rmat& Q = sb.Q;
    const rmat Qt = num::transpose(Q);
    const rmat A1sq = A1 * A1;
    const rmat W = Qt * A1sq * Q;
    std::vector<real> A1sq_din = num::matvec(A1sq, d_in);
    std::vector<real> tmp(K);
    for (unsigned int i = 0; i < K; i++) {
        tmp[i] = A1sq_din[i] - d_in[i];
    }
    const std::vector<real> cvec = num::matvec(Qt, tmp);
    const rmat MA1Q = M * A1 * Q;
    std::vector<real> avec(K, 0.0);
    for (unsigned int j = 0; j < K; j++) {
        real acc = 0.0;
        for (unsigned int i = 0; i < K; i++) {
            acc += v[i] * MA1Q(i, j);
        }
        avec[j] = M_CP * acc;
    }
    std::vector<real> A1din = num::matvec(A1, d_in);
    for (unsigned int i = 0; i < K; i++) {
        tmp[i] = A1din[i] - d_in[i];
    }
    const std::vector<real> Mt = num::matvec(M, tmp);
    real bsc = 0.0;
    for (unsigned int i = 0; i < K; i++) {
        bsc += v[i] * (Mt[i] + d_eq[i]);
    }
    bsc *= M_CP;
    const rmat Winv = num::solve_partial_pivot(W, rmat::identity(K));
    const rmat P = A1 * Q * Winv;
    const std::vector<real> Pc = num::matvec(P, cvec);
    std::vector<real> ovec(K);
    for (unsigned int i = 0; i < K; i++) {
        ovec[i] = (A1din[i] - d_in[i]) - Pc[i];
    }
    const rmat QtA1 = Qt * A1;

    std::vector<real>& z = *zblob;
    z[0] = 1.0;
    z[1] = (real)nth;
    z[2] = bsc;
    size_t q = 3;
    for (unsigned int j = 0; j < B; j++) {
        z[q + j] = (j < nth) ? sb.theta[j] * dt : 0.0;
    }
    q += B;
    for (unsigned int i = 0; i < K * K; i++) {
        z[q + i] = W.a[i];
    }
    q += K * K;
    for (unsigned int i = 0; i < K; i++) {
        z[q + i] = cvec[i];
        z[q + K + i] = avec[i];
    }
    q += 2 * K;
    for (unsigned int i = 0; i < K * K; i++) {
        z[q + i] = P.a[i];
    }
    q += K * K;
    for (unsigned int i = 0; i < K; i++) {
        z[q + i] = ovec[i];
    }
    q += K;
    for (unsigned int i = 0; i < K * K; i++) {
        z[q + i] = Qt.a[i];
        z[q + K * K + i] = QtA1.a[i];
    }
    q += 2 * K * K;
    for (unsigned int i = 0; i < K; i++) {
        z[q + i] = d_in[i];
    }
}

std::vector<real>
initial_density(density_algo algo, unsigned int N, const qm_operator* rho)
{
    switch (algo) {
    case density_algo::reg_cayley: {
        std::vector<real> d(2 * N * N, 0.0);
        if (rho) {
            const auto m = cv_representation::to_matrix(*rho, N);
            for (unsigned int i = 0; i < N * N; i++) {
                d[2 * i] = m.a[i].real();
                d[2 * i + 1] = m.a[i].imag();
            }
        }
        return d;
    }
    case density_algo::cvr_rodr:
        if (rho) {
            return cv_representation::coherence_vector(*rho, N);
        }
        return std::vector<real>(N * N - 1, 0.0);
    default:
        return std::vector<real>();
    }
}

void
fdtd_problem::bind()
{
    std::memset(&prob, 0, sizeof(prob));
    prob.abi_version = MBH_ABI_VERSION;
    prob.algo = (algo == density_algo::noop)
        ? MBH_ALGO_NOOP
        : (algo == density_algo::reg_cayley ? MBH_ALGO_REG_CAYLEY
                                            : MBH_ALGO_CVR_RODR);
    prob.num_levels = num_lvl;
    prob.num_materials = num_materials;
    prob.num_gridpoints = mat_idx.size();
    prob.num_timesteps = num_timesteps;
    prob.dx_inv = dx_inv;
    prob.fdtd_consts = fdtd_consts.data();
    prob.qm_consts = qm_consts.empty() ? nullptr : qm_consts.data();
    prob.qm_stride = qm_stride;
    prob.zform_consts = zform_consts.empty() ? nullptr : zform_consts.data();
    prob.zform_stride = zform_stride;
    prob.mat_idx = mat_idx.data();
    prob.e_init = e_init.data();
    prob.h_init = h_init.data();
    prob.d_init = d_init.empty() ? nullptr : d_init.data();
    prob.num_sources = sources.size();
    prob.sources = sources.empty() ? nullptr : sources.data();
    prob.source_data = source_data.empty() ? nullptr : source_data.data();
    prob.num_records = records.size();
    prob.records = records.empty() ? nullptr : records.data();
    prob.scratch_size = scratch_size;
}
}

namespace mbsolve {

void
build_fdtd_problem(std::shared_ptr<const device> dev,
                   std::shared_ptr<scenario> scen, density_algo algo,
                   unsigned int num_lvl, bool with_zform, fdtd_problem& out)
{
    if (!dev || !scen) {
        throw std::invalid_argument("Device or scenario is null.");
    }
    if (dev->get_regions().size() == 0) {
        throw std::invalid_argument("No regions in device!");
    }
    out = fdtd_problem();
    out.algo = algo;
    out.num_lvl = (algo == density_algo::noop) ? 0 : num_lvl;

    /* grid and time step with Courant number 0.5
     * (solver_cpu_fdtd.tpp:67) */
    init_fdtd_simulation(dev, scen, 0.5);
    const real dx = scen->get_gridpoint_size();
    const real dt = scen->get_timestep_size();
    const uint64_t nx = scen->get_num_gridpoints();
    const uint64_t nt = scen->get_num_timesteps();
    out.dx_inv = 1.0 / dx;
    out.num_timesteps = nt;

    /* per-material constants, indexed in the order of the used-material
     * set (solver_cpu_fdtd.tpp:73-89) */
    std::map<std::string, unsigned int> id_to_idx;
    unsigned int j = 0;
    for (const auto& id : dev->get_used_materials()) {
        auto mat = material::get_from_library(id);
        const sim_constants_fdtd f = get_fdtd_constants(dev, scen, mat);
        out.fdtd_consts.push_back(f.fac_a);
        out.fdtd_consts.push_back(f.fac_b);
        out.fdtd_consts.push_back(f.fac_c);
        out.fdtd_consts.push_back(f.gamma);
        if (algo == density_algo::reg_cayley) {
            const auto b = cayley_constants(mat->get_qm(), dt, num_lvl);
            out.qm_consts.insert(out.qm_consts.end(), b.begin(), b.end());
            out.qm_stride = b.size();
        } else if (algo == density_algo::cvr_rodr) {
            std::vector<real> b, z;
            rodr_constants(mat->get_qm(), dt, num_lvl, b,
                           with_zform ? &z : nullptr);
            out.qm_consts.insert(out.qm_consts.end(), b.begin(), b.end());
            out.qm_stride = b.size();
            if (with_zform) {
                out.zform_consts.insert(out.zform_consts.end(), z.begin(),
                                        z.end());
                out.zform_stride = z.size();
            }
        }
        out.material_ids.push_back(id);
        id_to_idx[id] = j++;
    }
    out.num_materials = j;

    /* per-cell material and initial state (solver_cpu_fdtd.tpp:113-154);
     * boundary points belong to the first matching region */
    const uint32_t dstride = (algo == density_algo::reg_cayley)
        ? 2 * num_lvl * num_lvl
        : (algo == density_algo::cvr_rodr ? num_lvl * num_lvl - 1 : 0);
    out.mat_idx.resize(nx);
    out.e_init.resize(nx);
    out.h_init.resize(nx + 1);
    out.d_init.assign(nx * dstride, 0.0);
    auto ic_d = scen->get_ic_density();
    auto ic_e = scen->get_ic_electric();
    auto ic_h = scen->get_ic_magnetic();
    if (!ic_e || !ic_h) {
        throw std::invalid_argument("Missing field initial conditions.");
    }
    for (uint64_t i = 0; i < nx; i++) {
        const real x = i * dx;
        int idx = -1;
        bool has_qm = false;
        for (const auto& reg : dev->get_regions()) {
            if ((x >= reg->get_x_start()) && (x <= reg->get_x_end())) {
                idx = id_to_idx[reg->get_material()->get_id()];
                has_qm = (bool)reg->get_material()->get_qm();
                break;
            }
        }
        if (idx < 0) {
            throw std::invalid_argument("region not found (grid point " +
                                        std::to_string(i) + ")");
        }
        out.mat_idx[i] = idx;
        if (dstride) {
            std::vector<real> d;
            if (has_qm) {
                if (!ic_d) {
                    throw std::invalid_argument(
                        "Missing density initial condition.");
                }
                const qm_operator rho = ic_d->initialize(x);
                d = initial_density(algo, num_lvl, &rho);
            } else {
                d = initial_density(algo, num_lvl, nullptr);
            }
            std::copy(d.begin(), d.end(), out.d_init.begin() + i * dstride);
        }
        out.e_init[i] = ic_e->initialize(x);
        out.h_init[i] = ic_h->initialize(x);
    }
    out.h_init[nx] = 0.0;

    /* records (solver_cpu_fdtd.tpp:157-182, copy_list_entry.hpp:127-171) */
    uint64_t scratch = 0;
    for (const auto& rec : scen->get_records()) {
        copy_list_entry e;
        std::memset(&e.dev, 0, sizeof(e.dev));
        if (rec->get_interval() <= dt) {
            e.dev.rows = nt;
            e.dev.interval_ratio = 1.0;
        } else {
            e.dev.rows = std::floor(scen->get_endtime() / rec->get_interval()) +
                1;
            e.dev.interval_ratio = dt / rec->get_interval();
        }
        if (rec->get_position() < 0.0) {
            e.dev.position_idx = 0;
            e.dev.cols = nx;
        } else {
            e.dev.position_idx = std::round(rec->get_position() / dx);
            e.dev.cols = 1;
            if (e.dev.position_idx >= nx) {
                throw std::invalid_argument("Record position outside the "
                                            "device: " + rec->get_name());
            }
        }
        e.res = std::make_shared<result>(rec->get_name(), e.dev.cols,
                                         e.dev.rows);
        e.dev.type = rec->get_type();
        e.dev.is_complex = rec->is_complex() ? 1 : 0;
        e.dev.offset_real = scratch;
        e.dev.offset_imag = scratch + e.dev.cols * e.dev.rows;
        e.dev.col_idx = rec->get_col();
        e.dev.row_idx = rec->get_row();
        if (rec->get_type() == record::type::density &&
            algo != density_algo::noop &&
            (e.dev.row_idx >= num_lvl || e.dev.col_idx >= num_lvl)) {
            throw std::invalid_argument("Requested result is not available: " +
                                        rec->get_name());
        }
        scratch += e.size();
        if (rec->is_complex()) {
            scratch += e.size();
        }
        out.records.push_back(e.dev);
        out.copy_list.push_back(e);
    }
    out.scratch_size = scratch;

    /* source table (solver_cpu_fdtd.tpp:184-202) */
    for (const auto& src : scen->get_sources()) {
        mbh_source s;
        s.type = (src->get_type() == source::hard_source) ? MBH_SRC_HARD
                                                          : MBH_SRC_SOFT;
        const real pos = src->get_position() / dx;
        if (!(pos >= 0.0) || pos >= (real)nx) {
            throw std::invalid_argument("Source position outside the "
                                        "device: " + src->get_name());
        }
        s.x_idx = (unsigned int)pos;
        out.sources.push_back(s);
        for (unsigned int n = 0; n < nt; n++) {
            out.source_data.push_back(src->get_value(n * dt));
        }
    }
    out.bind();
}
}
