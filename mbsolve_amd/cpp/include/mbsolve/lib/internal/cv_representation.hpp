/*
 * Coherence-vector representation of an N-level density matrix
 * (reference: mbsolve-lib/include/mbsolve/lib/internal/
 * coherence_vector_representation.hpp).
 *
 * rho = I / N + 1/2 sum_i d_i g_i with the K = N^2 - 1 generators of su(N)
 * in the reference order: N(N-1)/2 symmetric ("u", real parts),
 * N(N-1)/2 antisymmetric ("v", imaginary parts), N-1 diagonal ("w").
 */
#ifndef MBSOLVE_LIB_INTERNAL_CV_REPRESENTATION_H
#define MBSOLVE_LIB_INTERNAL_CV_REPRESENTATION_H

#include <memory>
#include <mbsolve/lib/internal/numerics.hpp>
#include <mbsolve/lib/qm_description.hpp>

namespace mbsolve {

class cv_representation
{
private:
    unsigned int m_num_levels;
    unsigned int m_num_adj;
    std::vector<num::cmat> m_generators;
    num::rmat m_hamiltonian;
    num::rmat m_dipole_op;
    std::vector<real> m_dipole_op_vec;
    num::rmat m_relax_superop;
    std::vector<real> m_equi;

    num::rmat liouvillian(const qm_operator& op) const;
    std::vector<real> op_vec(const qm_operator& op) const;

public:
    explicit cv_representation(std::shared_ptr<const qm_description> qm);

    /** Generators g_i (coherence_vector_representation.hpp:52-93). */
    const std::vector<num::cmat>& get_generators() const
    {
        return m_generators;
    }

    /** L_ij = Re(i/2 tr(H [g_i, g_j])) / hbar (:149-166). */
    const num::rmat& get_hamiltonian() const { return m_hamiltonian; }

    /** Liouvillian of the dipole operator. */
    const num::rmat& get_dipole_operator() const { return m_dipole_op; }

    /** v_i = Re tr(mu g_i) (:168-179). */
    const std::vector<real>& get_dipole_operator_vec() const
    {
        return m_dipole_op_vec;
    }

    /** G_ij = Re(1/2 tr(R(g_j) g_i)) (:181-197). */
    const num::rmat& get_relaxation_superop() const
    {
        return m_relax_superop;
    }

    /** d_eq_i = Re tr(R(I) g_i) / N (:199-213). */
    const std::vector<real>& get_equilibrium_vec() const { return m_equi; }

    /** Hermitian N x N matrix of a qm_operator (column-major upper
     * triangle of off-diagonal entries). */
    static num::cmat to_matrix(const qm_operator& op, unsigned int n);

    /** qm_operator of a matrix (diagonal real parts, upper triangle). */
    static qm_operator from_matrix(const num::cmat& m);

    /** Population rho_mm of a coherence vector (:289-336). */
    static real calc_population(const std::vector<real>& d, unsigned int n,
                                unsigned int m);

    /** Coherence vector of a density matrix, reproducing the reference
     * exactly (:341-394): the off-diagonal loops never advance their
     * output index, so all real parts land in d_0 (last one wins) and all
     * imaginary parts in d_{N(N-1)/2}.  Exact for diagonal rho and for
     * N = 2. */
    static std::vector<real> coherence_vector(const qm_operator& rho,
                                              unsigned int n);
};
}

#endif
