"""GPU feature coverage: batches of runs, source and record variants,
single-point (0-D) runs, all algorithm/level combinations, and a
physics check against the reference's digitized curves."""
import os

import numpy as np
import pytest

import mbsolve_amd.lib as mb
import mbsolve_amd.solverhip  # noqa: F401
from mbsolve_amd import capi, setups
import oracle

from conftest import rel_err

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))


def _solve(spec, form="", **opts):
    dev, scen = setups.build(spec)
    name = setups.solver_name(spec, form)
    print("case", spec["name"], name, opts, flush=True)
    sol = mb.create_hip_solver(name, dev, scen, **opts)
    sol.run()
    return sol


def test_batch_equals_single_runs(gpu):
    """Runs advanced together in one context give bit-identical records to
    the same runs done one by one (the pseudo cells decouple them)."""
    specs = setups.sit_2lvl_sweep(num_runs=5, min_log2=9, max_log2=11,
                                  endtime=0.3e-12, num_records=6)
    singles = []
    for s in specs:
        sol = _solve(s)
        singles.append([r.get_data_real() for r in sol.get_results()])
    mb.material.clear_library()
    runs = [setups.build(s, clear_library=False) for s in specs]
    batch = mb.solver_hip_batch("hip-fdtd-2lvl-reg-cayley", runs)
    batch.run()
    for r, ref in enumerate(singles):
        got = [x.get_data_real() for x in batch.get_results(r)]
        for a, b in zip(got, ref):
            assert np.array_equal(a, b), r


def test_batch_matches_oracle_mixed(gpu):
    specs = setups.sit_2lvl_sweep(num_runs=4, min_log2=9, max_log2=11,
                                  endtime=0.3e-12, num_records=6)
    mb.material.clear_library()
    runs = [setups.build(s, clear_library=False) for s in specs]
    batch = mb.solver_hip_batch("hip-fdtd-2lvl-reg-cayley-mixed", runs)
    batch.run()
    for r in range(len(specs)):
        ref = oracle.OracleRun(batch.problem(r)).run().records()
        got = [x.get_data_real() for x in batch.get_results(r)]
        for a, b in zip(got, ref):
            assert rel_err(a, b) < 1e-5


def _point_spec():
    """Soft + hard sources at interior cells, point records of every kind,
    density records of every population, three materials."""
    spec = setups.v3_rodr(num_gridpoints=1500, endtime=0.4e-12,
                          num_records=12)
    spec["materials"]["Glass"] = {"rel_permittivity": 2.25,
                                  "losses": 300.0,
                                  "overlap_factor": 0.8}
    spec["regions"].append(("glass", "Glass", 150e-6, 160e-6))
    spec["sources"].append({"kind": "gauss", "name": "g", "position": 50e-6,
                            "hard": False, "ampl": 3e7, "freq": 2.1e14,
                            "phase": 100e-15, "tau": 30e-15})
    spec["sources"].append({"kind": "sech", "name": "h", "position": 155e-6,
                            "hard": True, "ampl": 1e6, "freq": 1e14,
                            "phase": 5.0, "beta": 2e13})
    iv = 0.4e-12 / 40
    for name, t in (("e0", 0), ("p0", 1), ("h0", 2), ("i0", 4)):
        spec["records"].append({"name": name, "type": t, "row": 1, "col": 1,
                                "interval": iv, "position": 70e-6})
    for k in (1, 2, 3):
        spec["records"].append({"name": "d%d%d" % (k, k), "type": 3,
                                "row": k, "col": k, "interval": 0.0,
                                "position": 80e-6})
    spec["records"].append({"name": "d12", "type": 3, "row": 1, "col": 2,
                            "interval": iv, "position": 80e-6})
    return spec


@pytest.mark.parametrize("form,tol,dtol", [("", 1e-12, 1e-12),
                                           ("zform", 1e-9, 1e-10),
                                           ("mixed", 1e-5, 1e-6)])
def test_sources_records_materials(gpu, form, tol, dtol):
    """Field records: max|a-b|/max|b| <= tol.  Population records:
    max|a-b| <= dtol on the trace-1 scale of the density matrix (a small
    upper-level population has no scale of its own; FP32 storage of the
    density leaves ~1e-8 absolute there)."""
    spec = _point_spec()
    sol = _solve(spec, form)
    ref = oracle.OracleRun(sol.problem()).run().records()
    res = sol.get_results()
    assert len(res) == len(ref)
    for r, b in zip(res, ref):
        a = r.get_data_real()
        assert a.shape == b.shape
        if r.get_name().startswith("d"):
            err = float(np.max(np.abs(a - b)))
            assert err <= dtol, (r.get_name(), err)
        else:
            assert rel_err(a, b) <= tol, (r.get_name(), rel_err(a, b))
        if r.get_name() == "d12":
            assert not a.any()  # coherences are not recorded


@pytest.mark.parametrize("make", [
    lambda: setups.song2005(num_timesteps=4000),
    lambda: dict(setups.marskar2011(4, num_gridpoints=600, endtime=0.2e-12)),
    lambda: setups.marskar2011(5, num_gridpoints=300, endtime=0.08e-12),
    lambda: setups.v3_rodr(num_gridpoints=700, endtime=0.2e-12,
                           num_records=5, algo="reg-cayley"),
])
def test_more_levels_fp64(gpu, make):
    spec = make()
    sol = _solve(spec)
    ref = oracle.OracleRun(sol.problem()).run().records()
    for r, b in zip(sol.get_results(), ref):
        assert rel_err(r.get_data_real(), b) <= 1e-12, r.get_name()


@pytest.mark.parametrize("levels", [4, 5, 6])
def test_cayley_many_levels_fp64(gpu, levels):
    """Regularized Cayley for 4..6 levels (the reference's
    cpu-fdtd-<N>lvl-reg-cayley): the anharmonic ladder against the
    oracle's reference-order update."""
    spec = setups.marskar2011(levels, num_gridpoints=3000, endtime=0.2e-12)
    spec["algo"] = "reg-cayley"
    sol = _solve(spec)
    ref = oracle.OracleRun(sol.problem()).run().records()
    for r, b in zip(sol.get_results(), ref):
        assert rel_err(r.get_data_real(), b) <= 1e-12, r.get_name()


def test_noop_maxwell(gpu):
    spec = {"name": "vac", "algo": "noop", "num_levels": 0,
            "materials": {"Vacuum": {}, "Die": {"rel_permittivity": 4.0}},
            "regions": [("v", "Vacuum", 0.0, 60e-6),
                        ("d", "Die", 60e-6, 100e-6)],
            "num_gridpoints": 3001, "endtime": 0.4e-12,
            "rho_init": ([1.0, 0.0], []),
            "sources": [{"kind": "gauss", "name": "g", "position": 20e-6,
                         "hard": False, "ampl": 1e3, "freq": 3e13,
                         "phase": 60e-15, "tau": 20e-15}],
            "records": [{"name": "e", "interval": 0.4e-12 / 20},
                        {"name": "h", "interval": 0.4e-12 / 20}]}
    for form in ("", "mixed"):
        sol = _solve(spec, form)
        ref = oracle.OracleRun(sol.problem()).run().records()
        for r, b in zip(sol.get_results(), ref):
            tol = 0.0 if form == "" else 1e-5
            assert rel_err(r.get_data_real(), b) <= tol


@pytest.mark.parametrize("case,form,tile", [
    ("sit2", "", 256), ("sit2", "zform", 512), ("noop", "", 512),
    ("noop", "mixed", 512)])
def test_deepest_temporal_block(gpu, case, form, tile):
    """Temporal block 64 (the deepest, one bit per step in the record mask)
    and a request beyond it (clamped to 64): records every step, equal to
    the one-step-per-launch run (bitwise for the reference-order FP64 form,
    to the form's bound otherwise)."""
    if case == "sit2":
        spec = setups.sit_2lvl(num_gridpoints=2000, endtime=0.2e-12,
                               num_records=100000)
    else:
        spec = setups.maxwell_line(num_gridpoints=40000, num_timesteps=1200)
        spec["records"] = [{"name": "e", "interval": 0.0,
                            "position": 300e-9},
                           {"name": "h", "interval": 0.0,
                            "position": 400e-9}]
    base = _solve(spec, form, temporal_block=1, tile_cells=64)
    ref = [r.get_data_real() for r in base.get_results()]
    assert ref[0].shape[0] > 128
    for tb in (64, 100):
        sol = _solve(spec, form, temporal_block=tb, tile_cells=tile)
        assert sol.info()["temporal_block"] == 64
        for r, b in zip(sol.get_results(), ref):
            a = r.get_data_real()
            assert np.max(np.abs(b)) > 0.0
            if form == "":
                assert np.array_equal(a, b), (case, tb)
            else:
                tol = 1e-11 if form == "zform" else 1e-5
                assert rel_err(a, b) <= tol, (case, form, tb, rel_err(a, b))


def test_noop_mixed_held_h_start(gpu):
    """The FP32 stencil drops the factor that holds H at zero when every
    such cell starts at zero (its coefficient is zero there, so H stays
    zero); a state with H != 0 at those cells (h[0], h[N]) takes the kernel
    with the factor. Both match the oracle."""
    spec = setups.maxwell_line(num_gridpoints=2000, num_timesteps=600)
    spec["records"] = [{"name": "e", "interval": spec["endtime"] / 8}]
    dev, scen = setups.build(spec)
    for hz in (0.0, 0.5):
        prob = mb.build_problem(setups.solver_name(spec), dev, scen)
        h = np.array(prob["h_init"], dtype=float)
        h[0] = hz
        h[-1] = hz
        prob["h_init"] = h
        print("case held h", hz, flush=True)
        ctx = capi.Context([prob], capi.options("mixed", temporal_block=16))
        # the kernel without the H factor runs exactly when every held
        # cell starts at zero
        flags = ctx.info()["flags"]
        assert bool(flags & capi.MBH_INFO_HZ_ZERO) == (hz == 0.0), (hz, flags)
        ctx.run(0, int(prob["num_timesteps"]))
        got = ctx.records(0)
        orc = oracle.OracleRun(prob).run()
        want = orc.scratch[:len(got)]
        assert np.max(np.abs(want)) > 0
        assert rel_err(got, want) <= 1e-5, (hz, rel_err(got, want))
        # h[0] and h[N] are held at zero once step 0 has run, also in the
        # fetched state (the trailing pseudo cell is never written back by
        # a tile)
        st = ctx.state(0)
        ref = orc.state()
        assert st["h"][0] == 0.0 and st["h"][-1] == 0.0, hz
        assert ref["h"][0] == 0.0 and ref["h"][-1] == 0.0, hz
        assert rel_err(st["e"], ref["e"]) <= 1e-5, hz
        assert rel_err(st["h"], ref["h"]) <= 1e-5, hz
        ctx.close()


def test_state_after_partial_runs(gpu):
    """run_steps in pieces equals one call; fetch_state is the oracle's."""
    spec = setups.sit_2lvl(num_gridpoints=900, endtime=0.3e-12,
                           num_records=5)
    dev, scen = setups.build(spec)
    sol = mb.create_hip_solver("hip-fdtd-2lvl-reg-cayley", dev, scen,
                               temporal_block=7)
    nt = sol.problem()["num_timesteps"]
    for a, b in ((0, 3), (3, 100), (100, nt)):
        sol.run_steps(a, b)
    st = sol.fetch_state()
    run = oracle.OracleRun(sol.problem()).run()
    ref = run.state()
    for k in ("e", "h", "p", "d"):
        assert np.array_equal(st[k], ref[k]), k


def test_ziolkowski_sit_reference_curve(gpu):
    """Self-induced transparency set-up of the reference tool
    (mbsolve-tool.cpp:249-309) at its default 32768 cells to 200 fs; the
    inversion at t = 187.5 fs (row 75, as the reference's MATLAB script
    plots) against tools/reference-data/ziolkowski1995-2.csv, a digitized
    literature curve (loose features only)."""
    spec = setups.ziolkowski1995()
    sol = _solve(spec)
    inv = sol.get_results()[0].get_data_real()
    dx = 150e-6 / (32768 - 1)
    ref = np.loadtxt(os.path.join(HERE, "golden", "reference-data",
                                  "ziolkowski1995-2.csv"), delimiter=",")
    x = np.arange(inv.shape[1]) * dx
    sim = np.interp(ref[:, 0] * 1e-6, x, inv[75])
    # ahead of the pulse: ground state; pulse peak position and height
    assert np.all(sim[ref[:, 0] > 49.5] < -0.95)
    x_peak_ref = ref[np.argmax(ref[:, 1]), 0] * 1e-6
    win = (x > 30e-6) & (x < 50e-6)
    x_peak = x[win][np.argmax(inv[75][win])]
    assert abs(x_peak - x_peak_ref) < 1.0e-6
    assert inv[75][win].max() > 0.95


def test_hdf5_writer_after_run(gpu, tmp_path):
    """Results of a GPU run written by the hdf5 writer read back unchanged
    (writer-hdf5/src/writer_hdf5.cpp layout)."""
    hdf5 = pytest.importorskip("mbsolve_amd.writerhdf5")
    spec = setups.sit_2lvl(num_gridpoints=500, endtime=0.2e-12,
                           num_records=9)
    dev, scen = setups.build(spec)
    sol = mb.create_hip_solver("hip-fdtd-2lvl-reg-cayley", dev, scen)
    sol.run()
    fn = str(tmp_path / "sit.hdf")
    mb.writer.create_instance("hdf5").write(fn, sol.get_results(), dev,
                                            scen)
    f = hdf5.read_file(fn)
    assert f["attributes"]["timestep_size"] == scen.get_timestep_size()
    for r in sol.get_results():
        assert np.array_equal(f["results"][r.get_name()]["real"],
                              r.get_data_real())


@pytest.mark.parametrize("name", ["hip-fdtd-2lvl-reg-cayley",
                                  "hip-fdtd-3lvl-rodr-zform",
                                  "hip-fdtd-red-3lvl-rodr"])
def test_graph_replays_match_direct_launches(gpu, name):
    """Blocks of launches captured once in a hipGraph and replayed many
    times, with fetches (synchronisations, host work) between the runs:
    the state after every piece and the records equal those of direct
    launches exactly."""
    spec = (setups.sit_2lvl(num_gridpoints=1000, endtime=0.3e-12,
                            num_records=8)
            if "cayley" in name else
            setups.v3_rodr(num_gridpoints=1000, endtime=0.3e-12,
                           num_records=8))
    sols = []
    for graph in (1, 4):
        dev, scen = setups.build(spec)
        sols.append(mb.create_hip_solver(name, dev, scen, temporal_block=4,
                                         tile_cells=64,
                                         graph_launches=graph))
    assert sols[1].info()["graph_launches"] == 4
    nt = sols[0].problem()["num_timesteps"]
    bounds = list(range(0, nt, 50)) + [nt]
    for a, b in zip(bounds[:-1], bounds[1:]):
        for sol in sols:
            sol.run_steps(a, b)
        s0, s1 = sols[0].fetch_state(), sols[1].fetch_state()
        for key in ("e", "h", "p", "d"):
            assert np.array_equal(s0[key], s1[key]), (a, b, key)
    for r0, r1 in zip(sols[0].get_results(), sols[1].get_results()):
        assert np.array_equal(r0.get_data_real(), r1.get_data_real())


def test_red_order_batch_and_partial_runs(gpu):
    """The red step order in a batch of runs equals the runs done one by
    one, and run_steps in pieces equals one call; the final state is the
    oracle's red loop (three-level Rodrigues, FP64 form)."""
    dx = 150e-6 / 699  # one grid spacing, so all runs share dt and n_t
    specs = [setups.v3_rodr(num_gridpoints=n, endtime=0.2e-12, seed=s,
                            num_records=6, length=(n - 1) * dx,
                            name="v%d" % s)
             for s, n in ((1, 700), (2, 1000), (3, 450))]
    singles = []
    for sp in specs:
        dev, scen = setups.build(sp)
        name = setups.solver_name(sp, order="red")
        print("case red single", sp["name"], flush=True)
        sol = mb.create_hip_solver(name, dev, scen, temporal_block=5)
        nt = sol.problem()["num_timesteps"]
        for a, b in ((0, 7), (7, nt // 2), (nt // 2, nt)):
            sol.run_steps(a, b)
        sol.fetch_results()
        st = sol.fetch_state()
        ref = oracle.OracleRun(sol.problem()).run(order="red")
        for k, v in ref.state().items():
            assert rel_err(st[k], v) <= 1e-12, (sp["name"], k)
        singles.append([r.get_data_real() for r in sol.get_results()])
    mb.material.clear_library()
    runs = [setups.build(sp, clear_library=False) for sp in specs]
    print("case red batch", flush=True)
    batch = mb.solver_hip_batch("hip-fdtd-red-3lvl-rodr", runs)
    assert batch.info()["step_order"] == capi.MBH_ORDER_RED
    batch.run()
    for r, ref in enumerate(singles):
        for x, b in zip(batch.get_results(r), ref):
            assert np.array_equal(x.get_data_real(), b), r


@pytest.mark.parametrize("name,form", [
    ("hip-fdtd-2lvl-reg-cayley", "fp64"),
    ("hip-fdtd-2lvl-reg-cayley-mixed", "mixed"),
    ("hip-fdtd-3lvl-rodr-mfma", "mfma"),
    ("hip-fdtd-noop-mixed", "mixed"),
])
def test_reset_repeats_the_run(gpu, name, form):
    """mbh_reset rewinds to the initial state and clears the records: a
    second full run after it reproduces the first bit for bit (records and
    final state), also with a nonzero initial h[N] (read again at step 0
    of the repeated run)."""
    if "noop" in name:
        spec = setups.maxwell_line(num_gridpoints=3000, num_timesteps=700)
        spec["records"] = [{"name": "e", "interval": spec["endtime"] / 8}]
    elif "rodr" in name:
        spec = setups.v3_rodr(num_gridpoints=1500, endtime=0.3e-12,
                              num_records=8)
    else:
        spec = setups.sit_2lvl(num_gridpoints=1500, endtime=0.3e-12,
                               num_records=8)
    dev, scen = setups.build(spec)
    prob = mb.build_problem(name, dev, scen)
    if "noop" in name:
        h = np.array(prob["h_init"], dtype=float)
        h[-1] = 0.25
        prob["h_init"] = h
    print("case reset", name, flush=True)
    ctx = capi.Context([prob], capi.options(form))
    nt = int(prob["num_timesteps"])
    ctx.run(0, nt)
    rec1, st1 = ctx.records(0), ctx.state(0)
    assert np.max(np.abs(rec1)) > 0
    # the same launches again (FP32 storage rounds the state at every
    # launch boundary, so only the FP64 form is also independent of where
    # a run is split)
    pieces = [(0, nt // 3), (nt // 3, nt)] if form == "fp64" else [(0, nt)]
    for _ in range(2):
        ctx.reset()
        for a, b in pieces:
            ctx.run(a, b)
        rec2, st2 = ctx.records(0), ctx.state(0)
        assert np.array_equal(rec1, rec2)
        for k in ("e", "h", "p", "d"):
            assert np.array_equal(st1[k], st2[k]), k
    ctx.close()


def test_graph_launches_auto(gpu):
    """graph_launches = 0 groups the short launches of a small grid into
    captured graphs (even count); the records and the state equal those of
    launches one by one."""
    spec = setups.sit_2lvl(num_gridpoints=1200, endtime=0.6e-12,
                           num_records=16)
    out = []
    for graph in (1, 0):
        dev, scen = setups.build(spec)
        sol = mb.create_hip_solver("hip-fdtd-2lvl-reg-cayley-zform", dev,
                                   scen, graph_launches=graph)
        info = sol.info()
        print("case graph", graph, info, flush=True)
        if graph == 0:
            assert info["graph_launches"] >= 2
            assert info["graph_launches"] % 2 == 0
        sol.run()
        out.append(([r.get_data_real() for r in sol.get_results()],
                    sol.fetch_state(), info))
    # per-cell arithmetic does not depend on the tile geometry the
    # autotuner picks for either context
    (r1, s1, _), (r0, s0, _) = out
    for a, b in zip(r1, r0):
        assert np.array_equal(a, b)
    for k in ("e", "h", "p", "d"):
        assert np.array_equal(s1[k], s0[k]), k


@pytest.mark.parametrize("kind,form", [("sit2", "mixed"), ("sit2", "zform64"),
                                       ("maxwell", "mixed")])
def test_specialised_kernels_match_general(gpu, kind, form):
    """The kernels chosen from the constants (two-level Bloch form with a
    diagonal Hamiltonian and a real dipole; pure Maxwell with the held H
    cells starting at zero) drop only terms that are exact zeros: records
    and state equal those of the general kernels (MBH_OPT_GENERIC)."""
    if kind == "sit2":
        spec = setups.sit_2lvl(num_gridpoints=1500, endtime=0.4e-12,
                               num_records=12)
        flag = capi.MBH_INFO_BLOCH_REAL
    else:
        spec = setups.maxwell_line(num_gridpoints=3000, num_timesteps=700)
        spec["records"] = [{"name": "e", "interval": spec["endtime"] / 8}]
        flag = capi.MBH_INFO_HZ_ZERO
    dev, scen = setups.build(spec)
    suffix = "zform" if form == "zform64" else form
    prob = mb.build_problem(setups.solver_name(spec, suffix), dev, scen)
    nt = int(prob["num_timesteps"])
    out = []
    for generic in (False, True):
        print("case specialised", kind, form, "generic", generic, flush=True)
        ctx = capi.Context([prob], capi.options(form, temporal_block=8,
                                                generic=generic))
        flags = ctx.info()["flags"]
        assert bool(flags & flag) == (not generic), flags
        ctx.run(0, nt)
        out.append((ctx.records(0), ctx.state(0)))
        ctx.close()
    (r0, s0), (r1, s1) = out
    assert np.max(np.abs(r0)) > 0
    assert np.array_equal(r0, r1)
    for k in ("e", "h", "p", "d"):
        assert np.array_equal(s0[k], s1[k]), k


@pytest.mark.parametrize("form", ["mixed", "zform"])
def test_bloch_forms_independent_of_tile_width(gpu, form):
    """The two-level Bloch forms at 1, 2 and 4 cells per lane (64-, 128-
    and 256-cell tiles) and in tiles of 2 and 4 wavefronts (512 and 1024
    cells, inner edges exchanged through LDS) at one temporal
    block: the same records and state bit for bit (per-cell arithmetic does
    not depend on the tiling; with FP32 state the launch boundaries, fixed
    by the block, are the same)."""
    spec = setups.sit_2lvl(num_gridpoints=3000, endtime=0.4e-12,
                           num_records=12)
    out = []
    for tile in (64, 128, 256, 512, 1024):
        dev, scen = setups.build(spec)
        name = setups.solver_name(spec, form)
        sol = mb.create_hip_solver(name, dev, scen, temporal_block=16,
                                   tile_cells=tile)
        info = sol.info()
        print("case tile", form, tile, info, flush=True)
        assert info["tile_cells"] == tile
        sol.run()
        out.append(([r.get_data_real() for r in sol.get_results()],
                    sol.fetch_state()))
    for recs, st in out[1:]:
        for a, b in zip(recs, out[0][0]):
            assert np.array_equal(a, b)
        for k in ("e", "h", "p", "d"):
            assert np.array_equal(st[k], out[0][1][k]), k


@pytest.mark.parametrize("tile", [512, 1024])
def test_multi_wavefront_tiles_in_a_batch(gpu, tile):
    """Five two-level runs of uneven length in one context (run boundaries
    and the sources at each run's first cell fall inside tiles and at
    wavefront edges): tiles of several wavefronts give the records of the
    256-cell tiles bit for bit, FP64 state, temporal block 24."""
    dx = 150e-6 / 8191
    specs = [setups.sit_2lvl(num_gridpoints=n, endtime=0.3e-12,
                             num_records=8, length=(n - 1) * dx,
                             area=a, name="w%d" % i)
             for i, (n, a) in enumerate(((700, 6.0), (1300, 3.0), (257, 9.0),
                                         (2100, 6.3), (999, 4.0)))]
    out = []
    for tc in (256, tile):
        mb.material.clear_library()
        runs = [setups.build(sp, clear_library=False) for sp in specs]
        print("case batch tile", tc, flush=True)
        b = mb.solver_hip_batch("hip-fdtd-2lvl-reg-cayley-zform", runs,
                                temporal_block=24, tile_cells=tc)
        assert b.info()["tile_cells"] == tc
        b.run()
        out.append([[x.get_data_real() for x in b.get_results(r)]
                    for r in range(len(specs))])
    for r in range(len(specs)):
        for a, c in zip(out[0][r], out[1][r]):
            assert np.max(np.abs(a)) > 0
            assert np.array_equal(a, c), (tile, r)
