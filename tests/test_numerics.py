"""Host set-up numerics of the product against golden vectors (scipy
expm) and against the mathematics (skew-symmetric block structure,
linear solves)."""
import os

import numpy as np
import pytest

import mbsolve_amd.lib as mb

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = np.load(os.path.join(HERE, "golden", "golden_numerics.npz"))


def _pairs():
    keys = sorted(k for k in GOLDEN.files if k.startswith("A"))
    return [(GOLDEN[k], GOLDEN["E" + k[1:]]) for k in keys]


@pytest.mark.parametrize("idx", range(len(_pairs())))
def test_expm_matches_golden(idx):
    a, e = _pairs()[idx]
    got = mb.numerics.expm(a)
    scale = max(np.abs(e).max(), 1.0)
    assert np.abs(got - e).max() / scale < 2e-13 * max(1.0, np.abs(a).sum(0).max())


def test_expm_closed_forms():
    # diagonal and 2x2 upper-triangular (exact exponentials)
    d = np.diag([-1e-3, 0.2, -3.0])
    assert np.allclose(mb.numerics.expm(d), np.diag(np.exp(np.diag(d))),
                       rtol=1e-15, atol=0)
    a, b, c = -0.3, 0.7, -1.1
    t = np.array([[a, b], [0.0, c]])
    ref = np.array([[np.exp(a), b * (np.exp(a) - np.exp(c)) / (a - c)],
                    [0.0, np.exp(c)]])
    assert np.allclose(mb.numerics.expm(t), ref, rtol=1e-14, atol=1e-16)
    # rotation generator: exp(x [[0,-1],[1,0]])
    x = 0.9
    r = mb.numerics.expm(np.array([[0.0, -x], [x, 0.0]]))
    assert np.allclose(r, [[np.cos(x), -np.sin(x)], [np.sin(x), np.cos(x)]],
                       rtol=0, atol=2e-16)


@pytest.mark.parametrize("n,seed", [(3, 1), (8, 2), (15, 3), (24, 4),
                                    (35, 5)])
def test_skew_decompose_invariants(n, seed):
    rng = np.random.default_rng(seed)
    a = rng.standard_normal((n, n))
    u = a - a.T
    q, th = mb.numerics.skew_decompose(u)
    assert np.allclose(q.T @ q, np.eye(n), atol=1e-13)
    t = q.T @ u @ q
    blk = np.zeros((n, n))
    for j, tj in enumerate(th):
        blk[2 * j + 1, 2 * j] = tj
        blk[2 * j, 2 * j + 1] = -tj
    assert np.abs(t - blk).max() < 1e-12 * np.abs(u).max()
    assert np.all(np.diff(th) <= 0) and np.all(th > 0)
    # eigenvalues of U are +-i theta (and 0 for odd n)
    ev = np.sort(np.abs(np.linalg.eigvals(u).imag))[::-1]
    assert np.allclose(np.repeat(th, 2), ev[:2 * len(th)], rtol=1e-12)


def test_skew_decompose_degenerate():
    # two identical planes and a null space
    u = np.zeros((5, 5))
    u[1, 0], u[0, 1] = 2.0, -2.0
    u[3, 2], u[2, 3] = 2.0, -2.0
    rng = np.random.default_rng(9)
    rot = np.linalg.qr(rng.standard_normal((5, 5)))[0]
    u = rot @ u @ rot.T
    q, th = mb.numerics.skew_decompose(u)
    assert len(th) == 2 and np.allclose(th, [2.0, 2.0], rtol=1e-12)
    blk = np.zeros((5, 5))
    blk[:4, :4] = np.kron(np.eye(2), [[0.0, -2.0], [2.0, 0.0]])
    assert np.abs(q.T @ u @ q - blk).max() < 1e-12
    x = 0.37
    a2 = np.eye(5)
    for j in range(2):
        q1, q2 = q[:, 2 * j], q[:, 2 * j + 1]
        c1 = np.outer(q2, q1) - np.outer(q1, q2)
        c2 = -(np.outer(q1, q1) + np.outer(q2, q2))
        a2 += np.sin(th[j] * x) * c1 + (1 - np.cos(th[j] * x)) * c2
    import scipy.linalg
    assert np.allclose(a2, scipy.linalg.expm(x * u), atol=1e-14)


def test_full_pivot_solve():
    rng = np.random.default_rng(3)
    a = rng.standard_normal((12, 12))
    b = rng.standard_normal(12)
    x = mb.numerics.solve_full_pivot(a, list(b))
    assert np.allclose(a @ x, b, atol=1e-12)
