/*
 * Host set-up of an FDTD Maxwell-Bloch run: grid and time step, per-material
 * FDTD coefficients, per-material QM constant blobs in the layouts of
 * mbsolve_hip.h, initial density, source tables and record descriptors.
 * Together these perform what the reference solver constructor does
 * (solver-cpu/include/mbsolve/solver-cpu/solver_cpu_fdtd.tpp:36-203).
 */
#ifndef MBSOLVE_LIB_INTERNAL_FDTD_SETUP_H
#define MBSOLVE_LIB_INTERNAL_FDTD_SETUP_H

#include <memory>
#include <string>
#include <vector>
#include <mbsolve/lib/device.hpp>
#include <mbsolve/lib/result.hpp>
#include <mbsolve/lib/scenario.hpp>
#include <mbsolve_hip.h>

namespace mbsolve {

/** Grid point size, time step size and step count
 * (internal/common_fdtd.hpp:52-123). */
void init_fdtd_simulation(std::shared_ptr<const device> dev,
                          std::shared_ptr<scenario> scen, real courant = 1);

/** fac_a, fac_b, fac_c, gamma (internal/common_fdtd.hpp:131-190). */
struct sim_constants_fdtd {
    real fac_a;
    real fac_b;
    real fac_c;
    real gamma;
};
sim_constants_fdtd get_fdtd_constants(std::shared_ptr<const device> dev,
                                      std::shared_ptr<const scenario> scen,
                                      std::shared_ptr<const material> mat);

/** Density algorithm selector (the reference's template parameter). */
enum class density_algo { noop, reg_cayley, cvr_rodr };

/** Blob of lindblad_reg_cayley<N>::get_qm_constants
 * (algo_lindblad_reg_cayley.hpp:123-203), MBH_CAYLEY_STRIDE(N) doubles. */
std::vector<real> cayley_constants(std::shared_ptr<const qm_description> qm,
                                   real dt, unsigned int num_lvl);

/** Blob of lindblad_cvr_rodr<N>::get_qm_constants
 * (algo_lindblad_cvr_rodr.hpp:143-229), MBH_RODR_STRIDE(N) doubles, and
 * (optionally) the rotated-basis blob, MBH_ZFORM_STRIDE(N) doubles. */
void rodr_constants(std::shared_ptr<const qm_description> qm, real dt,
                    unsigned int num_lvl, std::vector<real>& blob,
                    std::vector<real>* zblob);

/** Initial density of one cell in the layout of the algorithm. */
std::vector<real> initial_density(density_algo algo, unsigned int num_lvl,
                                  const qm_operator* rho);

/** Per-record host bookkeeping (internal/copy_list_entry.hpp:113-229). */
struct copy_list_entry {
    mbh_record dev;
    std::shared_ptr<result> res;
    uint64_t size() const { return dev.rows * dev.cols; }
};

/** Everything mbh_create needs for one run; the vectors own the data the
 * mbh_problem points to. */
struct fdtd_problem {
    density_algo algo;
    unsigned int num_lvl;
    std::vector<real> fdtd_consts;
    std::vector<real> qm_consts;
    uint64_t qm_stride = 0;
    std::vector<real> zform_consts;
    uint64_t zform_stride = 0;
    std::vector<uint32_t> mat_idx;
    std::vector<real> e_init, h_init, d_init;
    std::vector<mbh_source> sources;
    std::vector<real> source_data;
    std::vector<copy_list_entry> copy_list;
    std::vector<mbh_record> records;
    uint64_t scratch_size = 0;
    uint32_t num_materials = 0;
    std::vector<std::string> material_ids;
    real dx_inv = 0.0;
    uint64_t num_timesteps = 0;
    mbh_problem prob;

    /** Refreshes `prob` so that it points into this object's vectors. */
    void bind();
};

/** Performs the reference constructor's set-up for (dev, scen) and the
 * given algorithm; mutates the scenario (dx, dt, n_t) like the
 * reference. with_zform also derives the rotated-basis Rodrigues blob. */
void build_fdtd_problem(std::shared_ptr<const device> dev,
                        std::shared_ptr<scenario> scen, density_algo algo,
                        unsigned int num_lvl, bool with_zform,
                        fdtd_problem& out);
}

#endif
