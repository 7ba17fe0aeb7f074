"""The oracle's red time loop (solver_cpu_fdtd_red.tpp:351-567) pinned on the
CPU: its results do not depend on the thread count or the overlap OL (the
redundant halo recomputation is exact), and its density-first step order is
the plain order (already pinned by the other oracle tests) started from the
density one update later:

    red(d0, p = 0)  ==  plain(d0' = D(d0, e0), p0' = p(d0'))

bit for bit, with E records equal at every step and the inversion of the red
run at step n equal to that of the plain run at step n - 1 (the one-step
inversion shift of the two reference solvers)."""
import numpy as np
import pytest

import mbsolve_amd.lib as mb
from mbsolve_amd import setups
import oracle


def _cases():
    return {
        "sit2": lambda nrec: setups.sit_2lvl(num_gridpoints=600,
                                             endtime=0.3e-12, phase=2.0,
                                             num_records=nrec),
        "v3": lambda nrec: setups.v3_rodr(num_gridpoints=500,
                                          endtime=0.25e-12,
                                          num_records=nrec),
        "qcl6": lambda nrec: setups.qcl6(num_gridpoints=120,
                                         num_timesteps=60,
                                         num_records=nrec),
        "noop": lambda nrec: setups.maxwell_line(num_gridpoints=3000,
                                                 num_timesteps=300),
    }


def _problem(case, nrec=16):
    spec = _cases()[case](nrec)
    dev, scen = setups.build(spec)
    return mb.build_problem(setups.solver_name(spec), dev, scen)


def _run(prob, **kw):
    r = oracle.OracleRun(prob).run(**kw)
    return r.records(), r.state()


@pytest.mark.parametrize("case", ["sit2", "v3", "qcl6", "noop"])
def test_red_independent_of_threads_and_overlap(case):
    prob = _problem(case)
    base_recs, base_state = _run(prob, order="red", threads=1)
    for threads, overlap in [(3, 0), (7, 0), (4, 3), (5, 1)]:
        recs, state = _run(prob, order="red", threads=threads,
                           overlap=overlap)
        for a, b in zip(recs, base_recs):
            assert np.array_equal(a, b), (case, threads, overlap)
        for k in ("e", "h", "p", "d"):
            assert np.array_equal(state[k], base_state[k]), (case, k)


def test_red_partial_runs_continue():
    prob = _problem("v3")
    n_t = int(prob["num_timesteps"])
    full_recs, full_state = _run(prob, order="red", threads=3)
    r = oracle.OracleRun(prob)
    r.run(0, n_t // 3, threads=3, order="red")
    r.run(n_t // 3, n_t, threads=2, order="red", overlap=5)
    for a, b in zip(r.records(), full_recs):
        assert np.array_equal(a, b)
    assert np.array_equal(r.state()["d"], full_state["d"])


def _advance_density(prob, run):
    """d0' = D(d0, e0), p0' = p(d0') per cell, with the oracle's own
    single-cell update (the first density update of the red order)."""
    n = int(prob["num_levels"])
    qs = int(prob["qm_stride"])
    q_all = np.asarray(prob["qm_consts"], dtype=float)
    mat = np.asarray(prob["mat_idx"])
    ds = run.ds
    upd = oracle.cayley_update if prob["algo"] == 1 else oracle.rodr_update
    for i in range(run.nx):
        q = q_all[qs * mat[i]:qs * (mat[i] + 1)]
        d, p = upd(n, q, run.d[i * ds:(i + 1) * ds], run.e[i])
        run.d[i * ds:(i + 1) * ds] = d
        run.pp[i] = p


@pytest.mark.parametrize("case", ["sit2", "v3", "qcl6"])
def test_red_is_plain_with_density_one_update_ahead(case):
    # records at every step: an interval below dt gives one row per step
    prob = _problem(case, nrec=100000)
    n_t = int(prob["num_timesteps"])
    recs = prob["records"]
    assert [int(r["rows"]) for r in recs] == [n_t] * len(recs)
    names = [int(r["type"]) for r in recs]
    red = oracle.OracleRun(prob).run(order="red", threads=3).records()

    plain_run = oracle.OracleRun(prob)
    _advance_density(prob, plain_run)
    plain = plain_run.run(order="plain", threads=3).records()

    for t, a, b in zip(names, red, plain):
        if t == 0:  # E
            assert np.array_equal(a, b), case
        else:       # inversion / populations trail by one step
            assert np.array_equal(a[1:], b[:-1]), case
    # and the two orders are different schemes: the inversion is not the
    # plain run's at the same step
    inv = names.index(4)
    assert not np.array_equal(red[inv][1:], plain[inv][1:])


def test_red_noop_equals_plain():
    prob = _problem("noop")
    red = oracle.OracleRun(prob).run(order="red", threads=4).records()
    plain = oracle.OracleRun(prob).run(order="plain", threads=4).records()
    for a, b in zip(red, plain):
        assert np.array_equal(a, b)
