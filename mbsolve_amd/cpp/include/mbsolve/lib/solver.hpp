/*
 * Solver factory and base class
 * (reference: mbsolve-lib/include/mbsolve/lib/solver.hpp,
 * mbsolve-lib/src/solver.cpp).
 */
#ifndef MBSOLVE_LIB_SOLVER_H
#define MBSOLVE_LIB_SOLVER_H

#include <functional>
#include <map>
#include <memory>
#include <string>
#include <vector>
#include <mbsolve/lib/device.hpp>
#include <mbsolve/lib/result.hpp>
#include <mbsolve/lib/scenario.hpp>
#include <mbsolve/lib/types.hpp>

namespace mbsolve {

class solver
{
public:
    typedef std::function<std::shared_ptr<solver>(
        std::shared_ptr<const device> dev,
        std::shared_ptr<scenario> scen)>
        bootstrap_t;

private:
    static std::map<std::string, bootstrap_t>& bootstraps();

protected:
    solver(
        const std::string& name,
        std::shared_ptr<const device> dev,
        std::shared_ptr<scenario> scen);

    std::string m_name;
    std::shared_ptr<const device> m_device;
    std::shared_ptr<scenario> m_scenario;
    std::vector<std::shared_ptr<result> > m_results;

public:
    virtual ~solver();

    static std::vector<std::string> get_avail_solvers();

    /** Throws std::invalid_argument("Solver already registered.") for a
     * duplicate name. */
    static void register_bootstrap(const std::string& name, bootstrap_t b);

    template<typename T>
    static void register_solver(const std::string& name)
    {
        register_bootstrap(
            name,
            [](std::shared_ptr<const device> dev,
               std::shared_ptr<scenario> scen) {
                return std::make_shared<T>(dev, scen);
            });
    }

    /** Throws std::invalid_argument("Unknown solver " + name). */
    static std::shared_ptr<solver> create_instance(
        const std::string& name,
        std::shared_ptr<const device> dev,
        std::shared_ptr<scenario> scen);

    const std::string& get_name() const;

    const scenario& get_scenario() const;

    const device& get_device() const;

    virtual void run() const = 0;

    const std::vector<std::shared_ptr<result> >& get_results() const;
};
}

#endif
