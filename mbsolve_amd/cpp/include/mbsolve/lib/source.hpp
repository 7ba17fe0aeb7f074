/*
 * Field sources, evaluated on the host into a per-step value table
 * (reference: mbsolve-lib/include/mbsolve/lib/source.hpp).
 */
#ifndef MBSOLVE_LIB_SOURCE_H
#define MBSOLVE_LIB_SOURCE_H

#include <cmath>
#include <string>
#include <mbsolve/lib/types.hpp>

namespace mbsolve {

/** Source base class: value(t) = amplitude * calc_value(t). */
class source
{
public:
    enum type
    {
        hard_source,
        soft_source
    };

protected:
    std::string m_name;
    real m_position;
    type m_type;
    real m_ampl;
    real m_freq;
    real m_phase;

public:
    source(
        const std::string& name,
        real position,
        type source_type,
        real ampl,
        real freq,
        real phase = 0)
      : m_name(name), m_position(position), m_type(source_type),
        m_ampl(ampl), m_freq(freq), m_phase(phase)
    {}

    virtual ~source() {}

    /** Source value at time t (source.hpp:94-106). */
    real get_value(real t, real /* current_value */ = 0.0) const
    {
        return m_ampl * calc_value(t);
    }

    virtual real calc_value(real /* t */) const { return 0.0; }

    const std::string& get_name() const { return m_name; }

    real get_position() const { return m_position; }

    type get_type() const { return m_type; }
};

/** sech pulse: sin(2 pi f t - phi_1) / cosh(beta t - phi_2)
 * (source.hpp:151-195). */
class sech_pulse : public source
{
private:
    real m_beta;
    real m_phase_sin;

public:
    sech_pulse(
        const std::string& name,
        real position,
        type source_type,
        real ampl,
        real freq,
        real phase,
        real beta,
        real phase_sin = 0.0)
      : source(name, position, source_type, ampl, freq, phase), m_beta(beta),
        m_phase_sin(phase_sin)
    {}

    real calc_value(real t) const
    {
        return 1 / std::cosh(m_beta * t - m_phase) *
            std::sin(2 * PI * m_freq * t - m_phase_sin);
    }
};

/** Gaussian pulse: exp(-(t - t0)^2 / tau^2) sin(2 pi f t)
 * (source.hpp:226-261). */
class gaussian_pulse : public source
{
private:
    real m_tau;

public:
    gaussian_pulse(
        const std::string& name,
        real position,
        type source_type,
        real ampl,
        real freq,
        real phase,
        real tau)
      : source(name, position, source_type, ampl, freq, phase), m_tau(tau)
    {}

    real calc_value(real t) const
    {
        return std::exp(-(t - m_phase) * (t - m_phase) / (m_tau * m_tau)) *
            std::sin(2 * PI * m_freq * t);
    }
};

/** Continuous wave: sin(2 pi f t + phase).  The reference declares such a
 * source only in a comment (source.hpp:129-145); it is needed for the weak
 * CW seed of the six-level gain-medium setups. */
class cw_source : public source
{
public:
    cw_source(
        const std::string& name,
        real position,
        type source_type,
        real ampl,
        real freq,
        real phase = 0.0)
      : source(name, position, source_type, ampl, freq, phase)
    {}

    real calc_value(real t) const
    {
        return std::sin(2 * PI * m_freq * t + m_phase);
    }
};
}

#endif
