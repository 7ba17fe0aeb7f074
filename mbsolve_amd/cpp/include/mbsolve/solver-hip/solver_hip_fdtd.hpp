/*
 * MI355X (gfx950) FDTD Maxwell-Bloch solvers behind the mbsolve solver
 * factory.  The constructor performs the reference constructor's set-up on
 * the host (solver-cpu/include/mbsolve/solver-cpu/solver_cpu_fdtd.tpp:36-203)
 * and hands the packed problem to the C ABI of mbsolve_hip.h; run() drives
 * the device time loop and copies the records into the results
 * (solver_cpu_fdtd.tpp:223-338).
 */
#ifndef MBSOLVE_SOLVER_HIP_FDTD_H
#define MBSOLVE_SOLVER_HIP_FDTD_H

#include <memory>
#include <string>
#include <utility>
#include <vector>
#include <mbsolve/lib/internal/fdtd_setup.hpp>
#include <mbsolve/lib/solver.hpp>
#include <mbsolve_hip.h>

namespace mbsolve {

/** Parses "hip-fdtd-noop[-mixed]",
 * "hip-fdtd-<N>lvl-rodr[-mixed|-zform|-mfma|-mfma64]" (MFMA forms for
 * N >= 3), "hip-fdtd-2lvl-reg-cayley[-mixed|-zform]",
 * "hip-fdtd-<N>lvl-reg-cayley[-mixed]" (N = 3..6) and the red step order
 * "hip-fdtd-red-noop", "hip-fdtd-red-<N>lvl-rodr",
 * "hip-fdtd-red-<N>lvl-reg-cayley" (FP64 form); returns false if not a
 * HIP name. */
bool parse_hip_solver_name(const std::string& name, density_algo& algo,
                           unsigned int& num_lvl, uint32_t& form);

/** MBH_ORDER_RED for "hip-fdtd-red-*" names, else MBH_ORDER_PLAIN. */
uint32_t hip_solver_order(const std::string& name);

/** Launch options from the environment: MBSOLVE_HIP_TEMPORAL_BLOCK,
 * MBSOLVE_HIP_TILE_CELLS, MBSOLVE_HIP_GRAPH_LAUNCHES, MBSOLVE_HIP_DEVICE. */
mbh_options hip_options_from_env(uint32_t form);

class solver_hip_fdtd : public solver
{
private:
    fdtd_problem m_problem;
    mbh_ctx* m_ctx = nullptr;
    uint32_t m_form;

public:
    solver_hip_fdtd(std::shared_ptr<const device> dev,
                    std::shared_ptr<scenario> scen, const std::string& name,
                    const mbh_options* opt = nullptr);

    ~solver_hip_fdtd();

    solver_hip_fdtd(const solver_hip_fdtd&) = delete;
    solver_hip_fdtd& operator=(const solver_hip_fdtd&) = delete;

    /** Runs all n_t steps and fills the results. */
    void run() const;

    /** Queues steps [n_begin, n_end) without copying records back. */
    void run_steps(uint64_t n_begin, uint64_t n_end) const;

    /** Rewinds to the initial state and clears the record scratch (queued;
     * the next run() repeats the simulation). */
    void reset() const;

    /** Waits for the device and copies the record scratch into the
     * results. */
    void fetch_results() const;

    /** Device time of the last run/run_steps in milliseconds. */
    double elapsed_ms() const;

    mbh_info info() const;

    uint32_t form() const { return m_form; }

    /** The packed problem handed to mbh_create (host copy). */
    const fdtd_problem& problem() const { return m_problem; }

    /** Current state (any pointer may be null), reference layouts. */
    void fetch_state(real* e, real* h, real* p, real* d) const;
};

/**
 * Several independent (device, scenario) runs advanced together by one
 * set of kernels (parameter sweeps).  All runs must use the same solver
 * name and end up with the same number of time steps.
 */
class solver_hip_batch
{
private:
    std::string m_name;
    std::vector<std::pair<std::shared_ptr<const device>,
                          std::shared_ptr<scenario> > >
        m_runs;
    std::vector<fdtd_problem> m_problems;
    mbh_ctx* m_ctx = nullptr;

public:
    solver_hip_batch(
        const std::string& name,
        const std::vector<std::pair<std::shared_ptr<const device>,
                                    std::shared_ptr<scenario> > >& runs,
        const mbh_options* opt = nullptr);

    ~solver_hip_batch();

    solver_hip_batch(const solver_hip_batch&) = delete;
    solver_hip_batch& operator=(const solver_hip_batch&) = delete;

    const std::string& get_name() const { return m_name; }

    size_t num_runs() const { return m_problems.size(); }

    void run() const;

    void run_steps(uint64_t n_begin, uint64_t n_end) const;

    /** Rewinds all runs to their initial state (see solver_hip_fdtd). */
    void reset() const;

    void fetch_results() const;

    double elapsed_ms() const;

    mbh_info info() const;

    /** Results of run r (filled by run()/fetch_results()). */
    std::vector<std::shared_ptr<result> > get_results(size_t r) const;

    const fdtd_problem& problem(size_t r) const { return m_problems[r]; }
};

/** Registers the hip-fdtd-* solver names with the factory (like
 * solver_cpu_loader, solver-cpu/src/solver_cpu_loader.cpp:32-89).
 * Constructing it twice is harmless. */
class solver_hip_loader
{
public:
    solver_hip_loader();
};
}

#endif
