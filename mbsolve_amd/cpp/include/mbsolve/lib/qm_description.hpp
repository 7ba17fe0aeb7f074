/*
 * Quantum-mechanical description of an active material
 * (reference: mbsolve-lib/include/mbsolve/lib/qm_description.hpp).
 */
#ifndef MBSOLVE_LIB_QM_DESCRIPTION_H
#define MBSOLVE_LIB_QM_DESCRIPTION_H

#include <memory>
#include <stdexcept>
#include <vector>
#include <mbsolve/lib/types.hpp>

namespace mbsolve {

/**
 * Hermitian operator given by its real main diagonal and the upper
 * off-diagonal triangle in column-major order:
 * (0,1), (0,2), (1,2), (0,3), (1,3), (2,3), ...
 * (reference qm_description.hpp:38-87).
 */
class qm_operator
{
private:
    unsigned int m_num_levels;
    std::vector<real> m_main_diag;
    std::vector<std::complex<real> > m_off_diag;

public:
    explicit qm_operator(
        const std::vector<real>& main_diagonal,
        const std::vector<std::complex<real> >& off_diagonal =
            std::vector<std::complex<real> >())
      : m_num_levels(main_diagonal.size()), m_main_diag(main_diagonal),
        m_off_diag(off_diagonal)
    {}

    virtual ~qm_operator() {}

    unsigned int get_num_levels() const { return m_num_levels; }

    const std::vector<real>& get_main_diagonal() const { return m_main_diag; }

    const std::vector<std::complex<real> >& get_off_diagonal() const
    {
        return m_off_diag;
    }
};

/** Superoperator acting on operators; zero by default
 * (reference qm_description.hpp:93-138). */
class qm_superop
{
protected:
    unsigned int m_num_levels;

public:
    explicit qm_superop(unsigned int num_levels) : m_num_levels(num_levels) {}

    virtual ~qm_superop() {}

    qm_operator operator()(const qm_operator& arg) const
    {
        return this->action(arg);
    }

    virtual qm_operator action(const qm_operator& /* arg */) const
    {
        return qm_operator(std::vector<real>(m_num_levels, 0.0));
    }

    unsigned int get_num_levels() const { return m_num_levels; }
};

/**
 * Lindblad relaxation: population scattering rates s[m][n] (from level n
 * to level m) and pure dephasing rates of the coherences
 * (reference qm_description.hpp:144-253).
 *
 * The constructor replaces s[m][m] by minus the total out-scattering rate
 * of level m and turns the pure dephasing rates into total (negative)
 * dephasing rates: -|pure_mn| - (|s_mm| + |s_nn|) / 2.
 */
class qm_lindblad_relaxation : public qm_superop
{
private:
    std::vector<std::vector<real> > m_scattering;
    std::vector<real> m_dephasing;

public:
    explicit qm_lindblad_relaxation(
        const std::vector<std::vector<real> >& scattering_rates,
        const std::vector<real>& pure_deph_rates);

    qm_operator action(const qm_operator& arg) const;

    const std::vector<std::vector<real> >& get_scattering_rates() const
    {
        return m_scattering;
    }

    const std::vector<real>& get_dephasing_rates() const
    {
        return m_dephasing;
    }
};

/** Quantum-mechanical description: carrier density, Hamiltonian, dipole
 * moment operator and relaxation superoperator
 * (reference qm_description.hpp:259-332). */
class qm_description
{
protected:
    unsigned int m_num_levels;
    real m_carrier_density;
    qm_operator m_hamiltonian;
    qm_operator m_dipole_op;
    std::shared_ptr<qm_superop> m_relax_superop;

public:
    explicit qm_description(
        real carrier_density,
        const qm_operator& hamiltonian,
        const qm_operator& dipole_operator,
        std::shared_ptr<qm_superop> relaxation_superop)
      : m_num_levels(hamiltonian.get_num_levels()),
        m_carrier_density(carrier_density), m_hamiltonian(hamiltonian),
        m_dipole_op(dipole_operator), m_relax_superop(relaxation_superop)
    {}

    virtual ~qm_description() {}

    unsigned int get_num_levels() const { return m_num_levels; }

    real get_carrier_density() const { return m_carrier_density; }

    const qm_operator& get_hamiltonian() const { return m_hamiltonian; }

    const qm_operator& get_dipole_operator() const { return m_dipole_op; }

    std::shared_ptr<qm_superop> get_relaxation_superop() const
    {
        return m_relax_superop;
    }
};

/**
 * Two-level system: H = diag(-hbar w0 / 2, +hbar w0 / 2),
 * mu_01 = -E0 d, scattering towards the equilibrium inversion and
 * dephasing (reference qm_description.hpp:338-432).
 */
class qm_desc_2lvl : public qm_description
{
private:
    real m_trans_freq;
    real m_dipole_mom;
    real m_scattering;
    real m_dephasing;
    real m_equi_inv;

public:
    explicit qm_desc_2lvl(
        real carrier_density,
        real transition_freq,
        real dipole_moment,
        real scattering_rate,
        real dephasing_rate,
        real equilibrium_inversion = -1.0);

    real get_transition_freq() const { return m_trans_freq; }

    real get_dipole_moment() const { return m_dipole_mom; }

    real get_scattering_rate() const { return m_scattering; }

    real get_dephasing_rate() const { return m_dephasing; }

    real get_equilibrium_inversion() const { return m_equi_inv; }
};
}

#endif
