"""Coherence-vector representation (product C++ and oracle numpy) against
the algebra of su(N) and brute-force superoperator actions."""
import numpy as np
import pytest

import mbsolve_amd.lib as mb
from mbsolve_amd import setups
from oracle import qm_setup


@pytest.mark.parametrize("n", [2, 3, 4, 5, 6])
def test_generators_orthonormal(n):
    g = qm_setup.generators(n)
    assert len(g) == n * n - 1
    for i, gi in enumerate(g):
        assert np.allclose(gi, gi.conj().T)
        assert abs(np.trace(gi)) < 1e-15
        for j, gj in enumerate(g):
            assert abs(np.trace(gi @ gj) - 2.0 * (i == j)) < 1e-14


def _qm_specs():
    return [setups.qcl6(num_gridpoints=64, num_timesteps=10)
            ["materials"]["AR_qcl6"]["qm"],
            setups.v3_rodr(num_gridpoints=64)["materials"]["AR_v3"]["qm"],
            setups.song2005()["materials"]["AR_Song"]["qm"],
            setups.marskar2011(4)["materials"]["AR_Marsk"]["qm"]]


@pytest.mark.parametrize("k", range(4))
def test_product_matrices_match_oracle(k):
    qs = _qm_specs()[k]
    dev, _ = setups.build({"name": "x", "algo": "cvr-rodr",
                           "num_levels": qs["num_levels"],
                           "materials": {"M": {"qm": qs}},
                           "regions": [("r", "M", 0.0, 1e-6)],
                           "num_gridpoints": 4, "endtime": 1e-15,
                           "rho_init": ([1.0] + [0.0] * (qs["num_levels"]
                                                       - 1), [])})
    qm = dev.get_regions()[0].get_material().get_qm()
    prod = mb.numerics.cv_matrices(qm)
    orc = qm_setup.cv_matrices(qs)
    for key in ("hamiltonian", "dipole_operator", "relaxation_superop"):
        a, b = prod[key], orc[key]
        assert np.abs(a - b).max() <= 1e-12 * max(np.abs(b).max(), 1e-300)
    for key in ("dipole_operator_vec", "equilibrium_vec"):
        a, b = prod[key], orc[key]
        assert np.abs(a - b).max() <= 1e-12 * max(np.abs(b).max(), 1e-300)
    # Liouvillians of Hermitian operators are real skew-symmetric
    L = prod["dipole_operator"]
    assert np.abs(L + L.T).max() <= 1e-13 * np.abs(L).max()


@pytest.mark.parametrize("k", range(4))
def test_relaxation_superop_brute_force(k):
    """d/dt d_i = sum_j G_ij d_j + d_eq_i reproduces the coherence vector of
    R(rho) for a random Hermitian trace-1 rho (rho = I/N + 1/2 d.g)."""
    q = qm_setup.qm_from_spec(_qm_specs()[k])
    n = q["num_levels"]
    cv = qm_setup.cv_matrices(q)
    g = qm_setup.generators(n)
    rng = np.random.default_rng(k)
    d = rng.standard_normal(n * n - 1) * 0.1
    rho = np.eye(n) / n + 0.5 * sum(di * gi for di, gi in zip(d, g))
    rr = qm_setup.relaxation_action(cv["s"], cv["deph"], rho)
    lhs = cv["relaxation_superop"] @ d + cv["equilibrium_vec"] * 1.0
    rhs = np.array([np.trace(rr @ gi).real for gi in g])
    assert np.allclose(lhs, rhs, rtol=1e-10, atol=1e-6 * np.abs(rhs).max())


@pytest.mark.parametrize("n", [2, 3, 4, 6])
def test_population_round_trip(n):
    rng = np.random.default_rng(n)
    pops = rng.uniform(size=n)
    pops /= pops.sum()
    d_prod = mb.numerics.coherence_vector(mb.qm_operator(list(pops)), n)
    d_orc = qm_setup.coherence_vector(list(pops), [], n)
    assert np.allclose(d_prod, d_orc, rtol=0, atol=1e-16)
    back = [mb.numerics.calc_population(list(d_prod), n, m)
            for m in range(n)]
    assert np.allclose(back, pops, atol=1e-15)
    assert abs(sum(back) - 1.0) < 1e-15
    import oracle
    lib_back = [qm_setup.cv_population(d_orc, n, m) for m in range(n)]
    assert np.allclose(lib_back, pops, atol=1e-15)


def test_coherence_vector_reference_behaviour():
    """Off-diagonal entries all land in d_0 (real parts) and d_{N(N-1)/2}
    (imaginary parts), last one winning, as in the reference."""
    op = mb.qm_operator([0.5, 0.3, 0.2], [0.1 + 0.2j, 0.05 - 0.01j, 0.03])
    d = mb.numerics.coherence_vector(op, 3)
    assert d[0] == 2 * 0.03 and d[3] == -2 * 0.0
    assert np.all(d[[1, 2, 4, 5]] == 0.0)
