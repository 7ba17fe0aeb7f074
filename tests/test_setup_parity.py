"""Product host set-up (C++, build_problem) against the oracle set-up
(numpy/scipy restatement), field by field, and the Rodrigues propagator
blobs against the matrix exponential they represent."""
import numpy as np
import pytest
import scipy.linalg

import mbsolve_amd.lib as mb
from mbsolve_amd import setups
from oracle import qm_setup

SPECS = {
    "sit2": lambda: setups.sit_2lvl(num_gridpoints=1500, endtime=0.4e-12,
                                    num_records=9),
    "v3": lambda: setups.v3_rodr(num_gridpoints=1200, endtime=0.3e-12,
                                 num_records=7),
    "qcl6": lambda: setups.qcl6(num_gridpoints=300, num_timesteps=120,
                                num_records=5),
    "zio": lambda: setups.ziolkowski1995(num_gridpoints=1000,
                                         endtime=20e-15),
    "song": lambda: setups.song2005(),
    "marskar4": lambda: setups.marskar2011(4, num_gridpoints=500,
                                           endtime=0.05e-12),
    "cpml": lambda: setups.tzenov2018_cpml(num_gridpoints=700,
                                           endtime=3e-12),
    # five levels, random (seeded) initial field, no source
    "tzenov": lambda: setups.tzenov2016(num_gridpoints=800, endtime=5e-12,
                                        seed=3),
}


def _both(spec, form=""):
    dev, scen = setups.build(spec)
    prod = mb.build_problem(setups.solver_name(spec, form), dev, scen)
    orc = qm_setup.build_problem(spec)
    return prod, orc


def _rel(a, b):
    a = np.asarray(a, float)
    b = np.asarray(b, float)
    s = np.abs(b).max() if b.size else 0.0
    d = np.abs(a - b).max() if a.size else 0.0
    return d / s if s > 0 else d


@pytest.mark.parametrize("name", sorted(SPECS))
def test_problem_fields(name):
    prod, orc = _both(SPECS[name]())
    for k in ("num_timesteps", "num_gridpoints", "num_materials",
              "scratch_size", "qm_stride", "algo", "num_levels"):
        assert prod[k] == orc[k], k
    assert prod["material_ids"] == orc["material_ids"]
    assert prod["dx_inv"] == orc["dx_inv"]
    for k in ("fdtd_consts", "e_init", "h_init", "d_init", "source_data"):
        assert _rel(prod[k], orc[k]) <= 1e-15, k
    assert np.array_equal(prod["mat_idx"], orc["mat_idx"])
    assert np.array_equal(prod["source_x"], orc["source_x"])
    assert np.array_equal(prod["source_type"], orc["source_type"])
    for rp, ro in zip(prod["records"], orc["records"]):
        for k in ("type", "row_idx", "col_idx", "rows", "cols",
                  "position_idx", "offset_real", "is_complex"):
            assert rp[k] == ro[k], k
        assert rp["interval_ratio"] == ro["interval_ratio"]


def _blobs(prod, orc, n):
    K = n * n - 1
    B = K // 2
    qs = prod["qm_stride"]
    for m in range(prod["num_materials"]):
        yield (prod["qm_consts"][m * qs:(m + 1) * qs],
               orc["qm_consts"][m * qs:(m + 1) * qs], K, B)


@pytest.mark.parametrize("name", ["sit2", "zio", "cpml"])
def test_cayley_constants(name):
    prod, orc = _both(SPECS[name]())
    assert prod["algo"] == 1
    assert _rel(prod["qm_consts"], orc["qm_consts"]) <= 1e-15
    n = prod["num_levels"]
    qs = prod["qm_stride"]
    for m in range(prod["num_materials"]):
        q = prod["qm_consts"][m * qs:(m + 1) * qs]
        if q[0] == 0:
            continue
        o = 5 + 4 * n * n
        d1 = q[o:o + n * n].reshape(n, n)
        D1 = q[o + 2 * n * n:o + 3 * n * n].reshape(n, n)
        dt = q[2]
        # D1 = exp(d1 dt / 2) (for two levels the closed form of a 2x2
        # matrix with zero column sums)
        assert np.allclose(D1, scipy.linalg.expm(d1 * dt / 2), rtol=1e-14,
                           atol=1e-18)
        assert np.allclose(D1.sum(axis=0), 1.0, atol=1e-15)


@pytest.mark.parametrize("name", ["v3", "qcl6", "song", "marskar4", "tzenov"])
def test_rodrigues_propagator(name):
    """A2(e) assembled from either blob equals expm(e dt U); A1 equals
    expm(M dt/2); M d_in = d_eq."""
    prod, orc = _both(SPECS[name]())
    n = prod["num_levels"]
    for qp, qo, K, B in _blobs(prod, orc, n):
        if qp[0] == 0:
            continue
        KK = K * K
        M = qp[5:5 + KK].reshape(K, K)
        U = qp[5 + KK:5 + 2 * KK].reshape(K, K)
        A1 = qp[5 + 3 * KK:5 + 4 * KK].reshape(K, K)
        o = 5 + 4 * KK
        d_eq = qp[o + K:o + 2 * K]
        d_in = qp[o + 2 * K:o + 3 * K]
        dt = qp[2]
        assert np.allclose(A1, scipy.linalg.expm(M * dt / 2), rtol=1e-13,
                           atol=1e-15)
        assert _rel(M, qo[5:5 + KK].reshape(K, K)) <= 1e-13
        assert _rel(A1, qo[5 + 3 * KK:5 + 4 * KK].reshape(K, K)) <= 1e-13
        if np.abs(d_eq).max() > 0:
            assert _rel(M @ d_in, d_eq) <= 1e-12
        for blob in (qp, qo):
            nth = int(blob[4])
            th = blob[o + 3 * K:o + 3 * K + B]
            C1 = blob[o + 3 * K + B:o + 3 * K + B + B * KK].reshape(B, K, K)
            C2 = blob[o + 3 * K + B + B * KK:].reshape(B, K, K)
            for e in (3.7e8, -1.2e9, 4e6):
                a2 = np.eye(K)
                for b in range(nth):
                    x = th[b] * e * dt
                    a2 = a2 + (np.sin(x) * C1[b] + (1 - np.cos(x)) * C2[b])
                ref = scipy.linalg.expm(e * dt * U)
                assert np.abs(a2 - ref).max() < 1e-12, (name, e)


@pytest.mark.parametrize("name", ["v3", "qcl6", "tzenov"])
def test_zform_blob_identities(name):
    """The rotated-basis blob reproduces the reference Strang step: from a
    random d, one reference step (A1, A2, A1) equals y -> W R y + c read out
    through P, o."""
    spec = SPECS[name]()
    dev, scen = setups.build(spec)
    prod = mb.build_problem(setups.solver_name(spec, "mixed"), dev, scen)
    n = prod["num_levels"]
    K = n * n - 1
    B = K // 2
    KK = K * K
    qs, zs = prod["qm_stride"], prod["zform_stride"]
    m = [i for i in range(prod["num_materials"])
         if prod["qm_consts"][i * qs] != 0][0]
    q = prod["qm_consts"][m * qs:(m + 1) * qs]
    z = prod["zform_consts"][m * zs:(m + 1) * zs]
    A1 = q[5 + 3 * KK:5 + 4 * KK].reshape(K, K)
    U = q[5 + KK:5 + 2 * KK].reshape(K, K)
    Mm = q[5:5 + KK].reshape(K, K)
    o = 5 + 4 * KK
    v, d_eq, d_in = q[o:o + K], q[o + K:o + 2 * K], q[o + 2 * K:o + 3 * K]
    dt, M_CP = q[2], q[1]
    th = z[3:3 + B]
    W = z[3 + B:3 + B + KK].reshape(K, K)
    c = z[3 + B + KK:3 + B + KK + K]
    a = z[3 + B + KK + K:3 + B + KK + 2 * K]
    P = z[3 + B + KK + 2 * K:3 + B + 2 * KK + 2 * K].reshape(K, K)
    oo = z[3 + B + 2 * KK + 2 * K:3 + B + 2 * KK + 3 * K]
    Qt = z[3 + B + 2 * KK + 3 * K:3 + B + 3 * KK + 3 * K].reshape(K, K)
    QtA1 = z[3 + B + 3 * KK + 3 * K:3 + B + 4 * KK + 3 * K].reshape(K, K)
    rng = np.random.default_rng(1)
    d = rng.standard_normal(K) * 0.3
    # a field that turns the fastest rotation by at most 0.5 rad per step
    # (at tens of radians the reference expm's own rounding dominates)
    e = min(2.3e8, 0.5 / np.abs(th).max())
    # reference step
    d1 = A1 @ (d + d_in) - d_in
    d2 = scipy.linalg.expm(e * dt * U) @ d1
    dref = A1 @ (d2 + d_in) - d_in
    pref = M_CP * v @ (Mm @ dref + d_eq)
    # z-form step
    y = QtA1 @ (d + d_in) - Qt @ d_in
    r = y.copy()
    for b in range(B):
        phi = th[b] * e
        cs, sn = np.cos(phi), np.sin(phi)
        r[2 * b] = cs * y[2 * b] - sn * y[2 * b + 1]
        r[2 * b + 1] = sn * y[2 * b] + cs * y[2 * b + 1]
    p = a @ r + z[2]
    y2 = W @ r + c
    assert _rel(P @ y2 + oo, dref) < 1e-12
    assert abs(p - pref) <= 1e-12 * abs(pref) + 1e-300
    # y* = -Q^T d_in is the fixed point at zero field, so the step can be
    # written y' = r + (W - I)(r - y*) (the MFMA kernels' form) and FP32
    # storage can hold y - y*
    yeq = -Qt @ d_in
    assert _rel(W @ yeq + c, yeq) < 1e-13
    assert _rel(r + (W - np.eye(K)) @ (r - yeq), y2) < 1e-12
    assert _rel(P @ yeq + oo, -d_in) < 1e-12
