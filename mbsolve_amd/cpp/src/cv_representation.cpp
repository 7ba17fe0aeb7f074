/*
 * Coherence-vector representation (see cv_representation.hpp).
 */
#include <cmath>
#include <stdexcept>
#include <mbsolve/lib/internal/cv_representation.hpp>

namespace mbsolve {

using num::cmat;
using num::rmat;

cmat
cv_representation::to_matrix(const qm_operator& op, unsigned int n)
{
    cmat m(n, n);
    for (unsigned int i = 0; i < n && i < op.get_main_diagonal().size();
         i++) {
        m(i, i) = op.get_main_diagonal()[i];
    }
    unsigned int row = 0, col = 1;
    for (size_t i = 0; i < op.get_off_diagonal().size() && col < n; i++) {
        m(row, col) = op.get_off_diagonal()[i];
        m(col, row) = std::conj(op.get_off_diagonal()[i]);
        if (row == col - 1) {
            row = 0;
            col++;
        } else {
            row++;
        }
    }
    return m;
}

qm_operator
cv_representation::from_matrix(const cmat& m)
{
    std::vector<real> diag;
    for (size_t i = 0; i < m.rows; i++) {
        diag.push_back(m(i, i).real());
    }
    std::vector<std::complex<real> > off;
    for (size_t col = 1; col < m.rows; col++) {
        for (size_t row = 0; row < col; row++) {
            off.push_back(m(row, col));
        }
    }
    return qm_operator(diag, off);
}

namespace {

std::complex<real> trace(const cmat& m)
{
    std::complex<real> t = 0.0;
    for (size_t i = 0; i < m.rows; i++) {
        t += m(i, i);
    }
    return t;
}

} /* namespace */

cv_representation::cv_representation(
    std::shared_ptr<const qm_description> qm)
  : m_num_levels(qm->get_num_levels()),
    m_num_adj(qm->get_num_levels() * qm->get_num_levels() - 1)
{
    const unsigned int N = m_num_levels;
    const std::complex<real> I(0.0, 1.0);

    /* generators: symmetric, antisymmetric, diagonal */
    for (unsigned int k = 0; k < N; k++) {
        for (unsigned int j = 0; j < k; j++) {
            cmat g(N, N);
            g(j, k) = 1.0;
            g(k, j) = 1.0;
            m_generators.push_back(g);
        }
    }
    for (unsigned int k = 0; k < N; k++) {
        for (unsigned int j = 0; j < k; j++) {
            cmat g(N, N);
            g(j, k) = std::complex<real>(0.0, -1.0);
            g(k, j) = std::complex<real>(0.0, +1.0);
            m_generators.push_back(g);
        }
    }
    for (unsigned int l = 0; l + 1 < N; l++) {
        cmat g(N, N);
        for (unsigned int j = 0; j <= l; j++) {
            g(j, j) = -1.0;
        }
        g(l + 1, l + 1) = (real)(l + 1);
        const real k = (real)(((l + 1) * (l + 2)) / 2);
        for (auto& x : g.a) {
            x /= std::sqrt(k);
        }
        m_generators.push_back(g);
    }

    m_hamiltonian = liouvillian(qm->get_hamiltonian());
    m_dipole_op = liouvillian(qm->get_dipole_operator());
    m_dipole_op_vec = op_vec(qm->get_dipole_operator());

    auto sop = qm->get_relaxation_superop();
    if (!sop) {
        throw std::invalid_argument("Missing relaxation superoperator.");
    }
    m_relax_superop = rmat(m_num_adj, m_num_adj);
    for (unsigned int j = 0; j < m_num_adj; j++) {
        const cmat rg = to_matrix((*sop)(from_matrix(m_generators[j])), N);
        for (unsigned int i = 0; i < m_num_adj; i++) {
            m_relax_superop(i, j) =
                (0.5 * trace(rg * m_generators[i])).real();
        }
    }
    const cmat ident = cmat::identity(N);
    const cmat rI = to_matrix((*sop)(from_matrix(ident)), N);
    m_equi.assign(m_num_adj, 0.0);
    for (unsigned int i = 0; i < m_num_adj; i++) {
        m_equi[i] = trace(rI * m_generators[i]).real() / N;
    }
}

rmat
cv_representation::liouvillian(const qm_operator& op) const
{
    const cmat o = to_matrix(op, m_num_levels);
    rmat L(m_num_adj, m_num_adj);
    for (unsigned int i = 0; i < m_num_adj; i++) {
        for (unsigned int j = 0; j < m_num_adj; j++) {
            const cmat& gi = m_generators[i];
            const cmat& gj = m_generators[j];
            const cmat comm = gi * gj - gj * gi;
            const std::complex<real> c =
                std::complex<real>(0.0, 1.0) * 0.5 * trace(o * comm);
            L(i, j) = c.real() / HBAR;
        }
    }
    return L;
}

std::vector<real>
cv_representation::op_vec(const qm_operator& op) const
{
    const cmat o = to_matrix(op, m_num_levels);
    std::vector<real> v(m_num_adj);
    for (unsigned int i = 0; i < m_num_adj; i++) {
        v[i] = trace(o * m_generators[i]).real();
    }
    return v;
}

real
cv_representation::calc_population(const std::vector<real>& d,
                                   unsigned int n, unsigned int m)
{
    if (n == 2) {
        return (m == 0) ? 0.5 - 0.5 * d[2] : 0.5 + 0.5 * d[2];
    }
    if (n == 3) {
        switch (m) {
        case 0:
            return 1.0 / 3 + 0.5 * (-1.0 * d[6] - 1.0 / std::sqrt(3.0) * d[7]);
        case 1:
            return 1.0 / 3 + 0.5 * (+1.0 * d[6] - 1.0 / std::sqrt(3.0) * d[7]);
        case 2:
            return 1.0 / 3 + 0.5 * (+2.0 / std::sqrt(3.0) * d[7]);
        default:
            return 0.0;
        }
    }
    real sum = 0.0;
    for (unsigned int i = 0; i + 1 < n; i++) {
        real factor;
        if (m < i + 1) {
            factor = -1.0;
        } else if (m == i + 1) {
            factor = (real)m;
        } else {
            factor = 0.0;
        }
        const int k = ((i + 1) * (i + 2)) / 2;
        sum += factor / std::sqrt((real)k) * d[n * (n - 1) + i];
    }
    return 1.0 / n + 0.5 * sum;
}

std::vector<real>
cv_representation::coherence_vector(const qm_operator& rho, unsigned int n)
{
    const unsigned int K = n * n - 1;
    std::vector<real> d(K, 0.0);
    const auto& off = rho.get_off_diagonal();
    /* the output index of the off-diagonal loops stays fixed, exactly as
     * in coherence_vector_representation.hpp:350-376 */
    for (size_t j = 0; j < off.size(); j++) {
        d[0] = 2 * off[j].real();
    }
    for (size_t j = 0; j < off.size(); j++) {
        d[(n * (n - 1)) / 2] = -2 * off[j].imag();
    }
    const auto& diag = rho.get_main_diagonal();
    for (unsigned int l = 0; l + 1 < n; l++) {
        real entry = 0.0;
        unsigned int j;
        for (j = 0; j <= l; j++) {
            entry += -1.0 * diag[j];
        }
        entry += (l + 1) * diag[j];
        const int k = ((l + 1) * (l + 2)) / 2;
        d[n * (n - 1) + l] = entry / std::sqrt((real)k);
    }
    return d;
}
}
