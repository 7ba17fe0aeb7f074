"""The oracle's boundary handling (solver_cpu_fdtd.tpp:238-275): the E
update of the last cell reads h[N] of the current state, and the boundary
condition sets h[0] = h[N] = 0 after every H update. Pinned by a locality
property on two runs that differ only in the initial h[N]: after one step
they differ in E only at the last cell and both hold h[N] = 0; after two
steps the difference has moved one cell further and no more (the stencil's
reach). The HIP kernels read the initial h[N] at step 0 only for the same
reason (the trailing pseudo cell is never written back)."""
import numpy as np

import mbsolve_amd.lib as mb
from mbsolve_amd import setups
import oracle


def _problem(hN, nx=60):
    dx = 1e-9
    spec = {"name": "line", "algo": "noop", "num_levels": 0,
            "materials": {"Die": {"rel_permittivity": 4.0, "losses": 10.0}},
            "regions": [("line", "Die", 0.0, dx * (nx - 1))],
            "num_gridpoints": nx, "endtime": 1e-15,
            "rho_init": ([1.0, 0.0], []),
            "records": [{"name": "e", "interval": 0.0, "position": 0.0}]}
    dev, scen = setups.build(spec)
    prob = mb.build_problem(setups.solver_name(spec), dev, scen)
    rng = np.random.default_rng(5)
    prob["e_init"] = rng.standard_normal(nx)
    h = rng.standard_normal(nx + 1)
    h[0] = 0.0
    h[nx] = hN
    prob["h_init"] = h
    return prob


def _after(prob, steps):
    r = oracle.OracleRun(prob)
    r.run(0, steps)
    return r.state()


def test_initial_h_at_the_end_is_read_once():
    nx = 60
    a0, b0 = _problem(0.0, nx), _problem(0.7, nx)
    assert int(a0["num_timesteps"]) >= 2
    for steps, reach in ((1, 1), (2, 2)):
        a, b = _after(a0, steps), _after(b0, steps)
        de = np.nonzero(a["e"] != b["e"])[0]
        dh = np.nonzero(a["h"] != b["h"])[0]
        assert a["h"][nx] == 0.0 and b["h"][nx] == 0.0
        assert a["h"][0] == 0.0 and b["h"][0] == 0.0
        # the difference starts at the last cell and spreads one cell per
        # step
        assert de.size > 0 and de.min() == nx - reach and de.max() == nx - 1
        assert dh.size > 0 and dh.min() >= nx - reach and dh.max() <= nx - 1
