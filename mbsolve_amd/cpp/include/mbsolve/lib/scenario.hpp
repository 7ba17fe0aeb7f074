/*
 * Simulation scenario: grid size, end time, initial conditions, sources
 * and records (reference: mbsolve-lib/include/mbsolve/lib/scenario.hpp,
 * mbsolve-lib/src/scenario.cpp).
 */
#ifndef MBSOLVE_LIB_SCENARIO_H
#define MBSOLVE_LIB_SCENARIO_H

#include <memory>
#include <random>
#include <string>
#include <vector>
#include <mbsolve/lib/qm_description.hpp>
#include <mbsolve/lib/record.hpp>
#include <mbsolve/lib/source.hpp>

namespace mbsolve {

/** Initial density matrix as a function of x. */
class ic_density
{
public:
    virtual ~ic_density() {}
    virtual qm_operator initialize(real x) = 0;
};

class ic_density_const : public ic_density
{
private:
    qm_operator m_rho;

public:
    explicit ic_density_const(const qm_operator& rho) : m_rho(rho) {}

    qm_operator initialize(real /* x */) { return m_rho; }
};

/** Initial field value as a function of x. */
class ic_field
{
public:
    virtual ~ic_field() {}
    virtual real initialize(real x) = 0;
};

class ic_field_const : public ic_field
{
private:
    real m_field_value;

public:
    explicit ic_field_const(real field_value) : m_field_value(field_value) {}

    real initialize(real /* x */) { return m_field_value; }
};

/** Normally distributed random field, amplitude * N(mean, stddev).  The
 * reference seeds it from std::random_device (scenario.hpp:125-157); the
 * optional seed makes runs reproducible. */
class ic_field_random : public ic_field
{
private:
    std::mt19937 m_rand_gen;
    std::normal_distribution<real> m_dis;
    real m_amplitude;

public:
    explicit ic_field_random(
        real mean = 0.0,
        real stddev = 1.0,
        real amplitude = 1e-15,
        unsigned int seed = 0)
      : m_rand_gen(seed ? seed : std::random_device()()), m_dis(mean, stddev),
        m_amplitude(amplitude)
    {}

    real initialize(real /* x */) { return m_dis(m_rand_gen) * m_amplitude; }
};

class scenario
{
private:
    std::string m_name;
    unsigned int m_num_timesteps;
    unsigned int m_num_gridpoints;
    real m_timestep_size;
    real m_gridpoint_size;
    real m_endtime;
    std::vector<std::shared_ptr<record> > m_records;
    std::vector<std::shared_ptr<source> > m_sources;
    std::shared_ptr<ic_density> m_dens_init;
    std::shared_ptr<ic_field> m_h_init;
    std::shared_ptr<ic_field> m_e_init;

public:
    scenario(
        const std::string& name,
        unsigned int num_gridpoints,
        real endtime,
        const qm_operator& rho_init,
        unsigned int num_timesteps = 2);

    scenario(
        const std::string& name,
        unsigned int num_gridpoints,
        real endtime,
        std::shared_ptr<ic_density> density_init,
        std::shared_ptr<ic_field> electric_init =
            std::make_shared<ic_field_random>(),
        std::shared_ptr<ic_field> magnetic_init =
            std::make_shared<ic_field_const>(0),
        unsigned int num_timesteps = 2);

    void add_record(std::shared_ptr<record> rec);

    const std::vector<std::shared_ptr<record> >& get_records() const;

    void add_source(std::shared_ptr<source> src);

    const std::vector<std::shared_ptr<source> >& get_sources() const;

    const std::string& get_name() const;

    unsigned int get_num_timesteps() const;

    void set_num_timesteps(unsigned int value);

    unsigned int get_num_gridpoints() const;

    void set_num_gridpoints(unsigned int value);

    real get_timestep_size() const;

    void set_timestep_size(real value);

    real get_gridpoint_size() const;

    void set_gridpoint_size(real value);

    real get_endtime() const;

    void set_endtime(real value);

    std::shared_ptr<ic_density> get_ic_density() const;

    void set_ic_density(std::shared_ptr<ic_density> density_init);

    std::shared_ptr<ic_field> get_ic_electric() const;

    void set_ic_electric(std::shared_ptr<ic_field> electric_init);

    std::shared_ptr<ic_field> get_ic_magnetic() const;

    void set_ic_magnetic(std::shared_ptr<ic_field> magnetic_init);
};
}

#endif
