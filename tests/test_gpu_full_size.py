"""GPU parity at the workloads' real sizes and against the reference's
digitized curves.

* The two-level SIT run at its full size (8192 cells, 1.5 ps = 49114
  steps): the FP64 form bit-identical to the oracle, the Bloch forms within
  their bounds.
* Sixteen runs of the 512-run sweep (its own seeded sizes, pulse areas and
  dephasing times; the runs below 11585 cells, so that the oracle finishes
  in seconds), FP32 storage and FP64, each run against its own oracle.
* song2005 (mbsolve-tool.cpp:163-248) and marskar2011 (mbsolve-tool.cpp:
  479-662) through the GPU forms against tools/reference-data.
* The six-level gain medium over a bench pass's 20000 steps (on 4096 of
  its cells) in the z-form and the MFMA form.
"""
import os

import numpy as np
import pytest

import mbsolve_amd.lib as mb
import mbsolve_amd.solverhip  # noqa: F401
from mbsolve_amd import setups
import oracle

from conftest import rel_err
from test_oracle_pins import check_marskar_curves, marskar_point_problem

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))
THREADS = min(16, os.cpu_count() or 1)


def _hip(spec, form, **opts):
    dev, scen = setups.build(spec)
    name = setups.solver_name(spec, form)
    print("case", spec["name"], name, flush=True)
    sol = mb.create_hip_solver(name, dev, scen, **opts)
    sol.run()
    return sol, [r.get_data_real() for r in sol.get_results()]


@pytest.fixture(scope="module")
def sit_full():
    """The oracle's records and final state of the full two-level SIT run
    (computed once for the module)."""
    spec = setups.sit_2lvl()
    dev, scen = setups.build(spec)
    prob = mb.build_problem(setups.solver_name(spec), dev, scen)
    assert int(prob["num_gridpoints"]) == 8192
    assert int(prob["num_timesteps"]) == 49114
    print("oracle sit2 full size", flush=True)
    run = oracle.OracleRun(prob).run(threads=THREADS)
    return spec, run.records(), run.state()


def test_sit_full_size_fp64_bitwise(gpu, sit_full):
    spec, ref, ref_state = sit_full
    sol, recs = _hip(spec, "")
    for a, b in zip(recs, ref):
        assert a.shape == b.shape == (256, 8192)
        assert np.array_equal(a, b)
    st = sol.fetch_state()
    for k in ("e", "h", "p", "d"):
        assert np.array_equal(st[k], ref_state[k]), k


@pytest.mark.parametrize("form,tol", [("zform", 1e-10), ("mixed", 1e-5)])
def test_sit_full_size_bloch_forms(gpu, sit_full, form, tol):
    spec, ref, _ = sit_full
    sol, recs = _hip(spec, form)
    for r, a, b in zip(sol.get_results(), recs, ref):
        err = rel_err(a, b)
        print("  ", form, r.get_name(), err, flush=True)
        assert 0.0 < err <= tol, (form, r.get_name(), err)


def _sweep_subset():
    specs = setups.sit_2lvl_sweep()
    small = [i for i, s in enumerate(specs) if s["num_gridpoints"] < 11585]
    rng = np.random.default_rng(5)
    pick = sorted(rng.choice(small, 16, replace=False).tolist())
    return [specs[i] for i in pick]


# FP32 storage rounds every run's state at each launch boundary; over the
# 49114 steps the error grows with the number of those roundings: at
# 16-step blocks the worst run of this subset reaches 1.1e-5 (at 32-step
# blocks of the 512-run batch, 5.7e-6 over all 512 runs).  The autotuner
# takes the deepest block within its trials' spread for FP32 storage, but
# an eight-run batch can still tune to 16, so the FP32 form is held to 2e-5
# here.
FP32_SWEEP_TOL = 2e-5


@pytest.mark.parametrize("half", [0, 1])
def test_sweep_runs_vs_oracle(gpu, half):
    """Eight of the sixteen runs per case as one batch in the FP64 form
    (bit-identical to each run's oracle), the Bloch form with FP64 storage
    (within 1e-10) and with FP32 storage (within FP32_SWEEP_TOL, and not
    equal: the FP32 rounding is there)."""
    specs = _sweep_subset()[8 * half:8 * (half + 1)]
    out = {}
    for name in ("hip-fdtd-2lvl-reg-cayley", "hip-fdtd-2lvl-reg-cayley-zform",
                 "hip-fdtd-2lvl-reg-cayley-mixed"):
        mb.material.clear_library()
        runs = [setups.build(s, clear_library=False) for s in specs]
        print("case sweep subset", half, name, flush=True)
        b = mb.solver_hip_batch(name, runs)
        print("  info", b.info(), flush=True)
        b.run()
        out[name] = ([[x.get_data_real() for x in b.get_results(r)]
                      for r in range(len(specs))],
                     [b.problem(r) for r in range(len(specs))])
    fp64, probs = out["hip-fdtd-2lvl-reg-cayley"]
    zform, _ = out["hip-fdtd-2lvl-reg-cayley-zform"]
    mixed, _ = out["hip-fdtd-2lvl-reg-cayley-mixed"]
    for r, prob in enumerate(probs):
        ref = oracle.OracleRun(prob).run(threads=THREADS).records()
        for a, z, m, o in zip(fp64[r], zform[r], mixed[r], ref):
            assert np.array_equal(a, o), r
            assert 0.0 < rel_err(z, o) <= 1e-10, (r, rel_err(z, o))
            err = rel_err(m, o)
            print("   run", r, "FP32 storage rel err", err, flush=True)
            assert 0.0 < err <= FP32_SWEEP_TOL, (r, err)


@pytest.mark.parametrize("form,tol", [("", 1e-12), ("zform", 1e-10),
                                      ("mixed", 1e-5), ("mfma64", 5e-6)])
def test_song2005_reference_curve_gpu(gpu, form, tol):
    """The single-point three-level V system against the digitized ground
    state population (tools/reference-data/song2005-3a.csv, features of a
    plot digitization) and against the oracle."""
    spec = setups.song2005()
    sol, recs = _hip(spec, form)
    ref = oracle.OracleRun(sol.problem()).run().records()
    assert float(np.max(np.abs(recs[0] - ref[0]))) <= tol
    d11 = recs[0][:, 0]
    dt = spec["endtime"] / (sol.problem()["num_timesteps"] - 1)
    csv = np.loadtxt(os.path.join(HERE, "golden", "reference-data",
                                  "song2005-3a.csv"), delimiter=",")
    t_ref = csv[:, 0] * 1e-15
    t = np.arange(len(d11)) * dt
    sim = np.interp(t_ref, t, d11)
    pre, post = t_ref < 38e-15, t_ref > 60e-15
    assert np.all(sim[pre] > 0.95)
    assert np.abs(sim[post] - csv[post, 1]).max() < 0.03
    win = (t > 40e-15) & (t < 52e-15)
    t_min = t[win][np.argmin(d11[win])]
    assert abs(t_min - t_ref[np.argmin(csv[:, 1])]) < 2e-15
    assert d11[win].min() < 0.2


@pytest.mark.parametrize("form,tol", [("", 1e-12), ("zform", 1e-11),
                                      ("mfma64", 1e-6), ("mfma", 1e-6)])
def test_marskar2011_reference_curves_gpu(gpu, form, tol):
    """marskar2011, six levels, with the tool's grid spacing (records at
    the source cell, which see only the source): rho55 and rho66 through
    the GPU forms against the oracle (populations, absolute) and the
    digitized curves of tools/reference-data."""
    spec = marskar_point_problem()
    sol, recs = _hip(spec, form)
    ref = oracle.OracleRun(sol.problem()).run().records()
    for r, a, b in zip(sol.get_results(), recs, ref):
        if r.get_name().startswith("d"):
            err = float(np.max(np.abs(a - b)))
        else:
            err = rel_err(a, b)
        assert err <= tol, (form, r.get_name(), err)
    dt = spec["endtime"] / (sol.problem()["num_timesteps"] - 1)
    check_marskar_curves(recs, dt)


@pytest.fixture(scope="module")
def six_level_pass():
    """The oracle's records of the six-level gain medium over the 20000
    steps of a bench pass on 4096 cells (computed once for the module)."""
    spec = setups.qcl6(num_gridpoints=4096, num_timesteps=20000,
                       num_records=16)
    dev, scen = setups.build(spec)
    prob = mb.build_problem(setups.solver_name(spec), dev, scen)
    print("oracle six-level 4096 x 20000", flush=True)
    return spec, oracle.OracleRun(prob).run(threads=THREADS).records()


@pytest.mark.parametrize("form,tol", [("zform", 1e-11), ("mfma64", 1e-6)])
def test_six_level_over_the_bench_pass_length(gpu, six_level_pass, form,
                                              tol):
    """The six-level gain medium over the 20000 steps of a bench pass on
    4096 cells (the workload's dx, so its dt) against the oracle: the
    FP64 z-form within 1e-11, the MFMA form within 1e-6 (inversion and
    populations absolute, E relative to its maximum)."""
    spec, ref = six_level_pass
    sol, recs = _hip(spec, form)
    for r, a, b in zip(sol.get_results(), recs, ref):
        if r.get_name().startswith(("d", "inv")):
            err = float(np.max(np.abs(a - b)))
        else:
            err = rel_err(a, b)
        print("  ", form, r.get_name(), err, flush=True)
        assert np.max(np.abs(b)) > 0
        assert err <= tol, (form, r.get_name(), err)
