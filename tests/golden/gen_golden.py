"""Writes the golden vectors of the set-up numerics to golden_numerics.npz.

The expected values come from scipy (a library, not the reference):
scipy.linalg.expm implements the scaling-and-squaring Pade method of
Higham (2005), the algorithm of Eigen's MatrixBase::exp used by the
reference set-up (algo_lindblad_reg_cayley.hpp:169,
algo_lindblad_cvr_rodr.hpp:185).  Inputs are seeded random matrices over the
norm ranges of each Pade degree (3/5/7/9/13 and squaring) plus the physical
matrices M dt/2 of the two- and three-level set-ups.

Run:  python tests/golden/gen_golden.py
"""
import os
import sys

import numpy as np
import scipy.linalg

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)


def main():
    rng = np.random.default_rng(20250101)
    mats = []
    for n in (2, 3, 8, 15, 35):
        for norm in (1e-3, 0.1, 0.5, 1.5, 4.0, 40.0, 700.0):
            a = rng.standard_normal((n, n))
            a *= norm / np.abs(a).sum(axis=0).max()
            mats.append(a)
    # skew-symmetric generators (Rodrigues propagators)
    for n in (3, 8, 15, 24, 35):
        a = rng.standard_normal((n, n))
        mats.append((a - a.T) * 0.7)
    # physical: M dt / 2 of the oracle set-up
    from oracle import qm_setup
    from mbsolve_amd import setups
    for spec in (setups.sit_2lvl(), setups.v3_rodr(num_gridpoints=4096),
                 setups.qcl6(num_gridpoints=4096, num_timesteps=100)):
        q = qm_setup.qm_from_spec(spec["materials"]["AR_" + spec["name"]]
                                  ["qm"])
        if q["num_levels"] == 2:
            s, _ = qm_setup.lindblad_relaxation(q["scattering"],
                                                q["pure_dephasing"])
            mats.append(s * 3.05e-17 / 2)
        cv = qm_setup.cv_matrices(q)
        M = cv["hamiltonian"] + cv["relaxation_superop"]
        mats.append(M * 3.05e-17 / 2)
    out = {}
    for i, a in enumerate(mats):
        out["A%03d" % i] = a
        out["E%03d" % i] = scipy.linalg.expm(a)
    np.savez_compressed(os.path.join(HERE, "golden_numerics.npz"), **out)
    print("wrote %d expm pairs" % len(mats))


if __name__ == "__main__":
    main()
