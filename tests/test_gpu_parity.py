"""GPU parity: the HIP solvers against the CPU FP64 oracle on identical
host-computed constants (same mbh_problem), through the solver factory and
directly through the C ABI."""
import numpy as np
import pytest

import mbsolve_amd.lib as mb
import mbsolve_amd.solverhip  # noqa: F401  (registers hip-fdtd-*)
from mbsolve_amd import capi, setups
import oracle

from conftest import rel_err

pytestmark = pytest.mark.gpu

FP64_TOL = 1e-12   # bound for the FP64 form against the oracle
MIXED_TOL = 1e-5   # bound for the FP32-storage form


def _run_hip(spec, form="", order="plain", **opts):
    dev, scen = setups.build(spec)
    name = setups.solver_name(spec, form, order)
    print("case", spec["name"], name, opts, flush=True)
    sol = mb.create_hip_solver(name, dev, scen, **opts)
    sol.run()
    return sol, [r.get_data_real() for r in sol.get_results()]


def _oracle_records(sol):
    return oracle.OracleRun(sol.problem()).run().records()


SMALL = [
    ("sit2", lambda: setups.sit_2lvl(num_gridpoints=2048, endtime=0.9e-12,
                                     num_records=32)),
    ("v3", lambda: setups.v3_rodr(num_gridpoints=2048, endtime=0.9e-12,
                                  num_records=32)),
    ("qcl6", lambda: setups.qcl6(num_gridpoints=512, num_timesteps=400,
                                 num_records=8)),
    ("zio", lambda: setups.ziolkowski1995(num_gridpoints=2048,
                                          endtime=60e-15)),
    ("tzenov", lambda: setups.tzenov2016(num_gridpoints=1024,
                                         endtime=20e-12)),
]


@pytest.mark.parametrize("case", [c[0] for c in SMALL])
def test_fp64_matches_oracle(gpu, case):
    spec = dict(SMALL)[case]()
    sol, recs = _run_hip(spec)
    ref = _oracle_records(sol)
    for a, b in zip(recs, ref):
        assert a.shape == b.shape
        assert rel_err(a, b) <= FP64_TOL, (case, rel_err(a, b))


@pytest.mark.parametrize("tb,tile", [(1, 64), (5, 128), (16, 256), (32, 256),
                                     (28, 64)])
def test_fp64_independent_of_tiling(gpu, tb, tile):
    spec = setups.sit_2lvl(num_gridpoints=1500, endtime=0.5e-12,
                           num_records=16)
    _, base = _run_hip(spec, temporal_block=1, tile_cells=64)
    _, recs = _run_hip(spec, temporal_block=tb, tile_cells=tile)
    for a, b in zip(recs, base):
        assert np.array_equal(a, b)


def _interfaces(case):
    if case == "gain_absorber":
        spec = setups.tzenov2018_cpml(num_gridpoints=1200, endtime=2e-12)
        iv = spec["endtime"] / 16
        spec["records"] = [{"name": "inv12", "interval": iv},
                           {"name": "e", "interval": iv},
                           {"name": "d22", "type": 3, "row": 2, "col": 2,
                            "interval": iv}]
        return spec
    if case == "sit2":
        return setups.sit_2lvl(num_gridpoints=1300, endtime=0.4e-12,
                               num_records=16)
    return setups.v3_rodr(num_gridpoints=1300, endtime=0.3e-12,
                          num_records=16)


@pytest.mark.parametrize("case,form,tiles,tol", [
    ("gain_absorber", "", (64, 128, 256), FP64_TOL),
    ("gain_absorber", "zform", (64, 128, 256, 512), 1e-11),
    ("sit2", "", (64, 128, 256), FP64_TOL),
    ("sit2", "zform", (64, 128, 256, 512), 1e-11),
    ("v3", "", (64, 128), FP64_TOL),
    ("v3", "zform", (64, 128), 1e-10),
])
def test_interface_tiles(gpu, case, form, tiles, tol):
    """Tiles at material interfaces: one density material beside vacuum
    (at two or more cells per lane some lanes straddle the interface) and
    two density materials in one tile (gain | absorber | gain). Every
    tiling matches the oracle; the FP64 form is bitwise the same for
    every tiling."""
    ref = base = None
    for tile in tiles:
        sol, recs = _run_hip(_interfaces(case), form, temporal_block=8,
                             tile_cells=tile)
        if ref is None:
            ref = _oracle_records(sol)
        for a, b in zip(recs, ref):
            assert rel_err(a, b) <= tol, (case, form, tile, rel_err(a, b))
        if form == "":
            if base is None:
                base = recs
            for a, b in zip(recs, base):
                assert np.array_equal(a, b), (case, tile)


def test_mixed_rodr_within_tolerance(gpu):
    spec = setups.v3_rodr(num_gridpoints=2048, endtime=0.9e-12,
                          num_records=32)
    sol, recs = _run_hip(spec, "mixed")
    ref = _oracle_records(sol)
    for a, b in zip(recs, ref):
        assert rel_err(a, b) <= MIXED_TOL, rel_err(a, b)


def test_zform64_close_to_reference_form(gpu):
    spec = setups.v3_rodr(num_gridpoints=2048, endtime=0.9e-12,
                          num_records=32)
    sol, recs = _run_hip(spec, "zform")
    ref = _oracle_records(sol)
    for a, b in zip(recs, ref):
        assert rel_err(a, b) <= 1e-10, rel_err(a, b)


@pytest.mark.parametrize("form,tol", [("zform", 1e-11), ("mixed", 1e-5)])
def test_cayley_bloch_forms(gpu, form, tol):
    """Two-level Cayley in closed Bloch form (FP64 / FP32 state) against
    the oracle's reference-order update: SIT pulse with finite T1/T2 and
    a soft source, every record kind."""
    spec = setups.sit_2lvl(num_gridpoints=3000, endtime=1.0e-12, t1=5e-13,
                           t2=2e-13, num_records=40)
    spec["records"] += [
        {"name": "d11", "type": 3, "row": 1, "col": 1, "interval": 0.0,
         "position": 75e-6},
        {"name": "d22", "type": 3, "row": 2, "col": 2, "interval": 0.0,
         "position": 75e-6},
        {"name": "p", "interval": 1e-12 / 39},
        {"name": "h", "interval": 1e-12 / 39}]
    sol, recs = _run_hip(spec, form)
    ref = _oracle_records(sol)
    for r, a, b in zip(sol.get_results(), recs, ref):
        if r.get_name().startswith("d"):
            err = float(np.max(np.abs(a - b)))
        else:
            err = rel_err(a, b)
        assert err <= tol, (r.get_name(), err)
    st = sol.fetch_state()
    rs = oracle.OracleRun(sol.problem()).run().state()
    scale = {"e": np.max(np.abs(rs["e"])), "h": np.max(np.abs(rs["h"])),
             "d": 1.0}
    for k in ("e", "h", "d"):
        assert np.max(np.abs(st[k] - rs[k])) <= tol * scale[k], k


def test_capi_direct(gpu):
    spec = setups.sit_2lvl(num_gridpoints=1000, endtime=0.3e-12,
                           num_records=8)
    dev, scen = setups.build(spec)
    prob = mb.build_problem("hip-fdtd-2lvl-reg-cayley", dev, scen)
    ctx = capi.Context([prob], capi.options("fp64", temporal_block=8))
    nt = int(prob["num_timesteps"])
    ctx.run(0, nt // 2)
    ctx.run(nt // 2, nt)
    scratch = ctx.records(0)
    ref = oracle.OracleRun(prob).run()
    assert np.array_equal(scratch, ref.scratch[:len(scratch)])
    st = ctx.state(0)
    rs = ref.state()
    for k in ("e", "h", "p", "d"):
        assert np.array_equal(st[k], rs[k]), k
    ctx.close()



def test_capi_device_allocation_failure(gpu):
    """A record scratch beyond the device's memory (2^36 doubles,
    512 GiB): mbh_create returns MBH_ERR_NOMEM (the reference checks every
    device allocation the same way, common_cuda.hpp:60-66) after releasing
    what it had allocated, and leaves no HIP error behind: the next context
    in the same process runs bit-identical to the oracle."""
    spec = setups.sit_2lvl(num_gridpoints=300, endtime=0.1e-12,
                           num_records=4)
    dev, scen = setups.build(spec)
    prob = mb.build_problem("hip-fdtd-2lvl-reg-cayley", dev, scen)
    big = dict(prob)
    big["scratch_size"] = 1 << 36
    print("case allocation failure", flush=True)
    with pytest.raises(capi.MBHError, match="hipMalloc") as ei:
        capi.Context([big], capi.options("fp64"))
    assert ei.value.code == 3  # MBH_ERR_NOMEM
    print("case run after the failure", flush=True)
    ctx = capi.Context([prob], capi.options("fp64"))
    ctx.run(0, int(prob["num_timesteps"]))
    scratch = ctx.records(0)
    ref = oracle.OracleRun(prob).run()
    assert np.any(scratch != 0.0)
    assert np.array_equal(scratch, ref.scratch[:len(scratch)])
    ctx.close()


MFMA_TOL = 1e-6  # split-FP16 generator product, short runs


def _qcl6_with_populations():
    spec = setups.qcl6(num_gridpoints=2000, num_timesteps=2500,
                       num_records=10, cw_ampl=3e6)
    for i in (1, 3, 6):
        spec["records"].append({"name": "d%d%d" % (i, i), "type": 3,
                                "row": i, "col": i, "interval": 0.0,
                                "position": 1.0e-6})
    return spec


@pytest.mark.parametrize("form", ["mfma", "mfma64"])
@pytest.mark.parametrize("make,tol", [
    (_qcl6_with_populations, MFMA_TOL),
    (lambda: setups.marskar2011(4, num_gridpoints=700, endtime=0.2e-12),
     MFMA_TOL),
    (lambda: setups.marskar2011(5, num_gridpoints=400, endtime=0.1e-12),
     MFMA_TOL),
    # three levels at optical frequencies, 7371 steps: W - I carries the
    # free precession (~5e-3 per step), so the split product's per-step
    # rounding accumulates faster
    (lambda: setups.v3_rodr(num_gridpoints=2048, endtime=0.9e-12,
                            num_records=32), 5e-6),
    # vacuum margins: uniform tiles without a density record populations
    # 1/N like the reference's zero coherence vector
    (lambda: _v3_three_regions(), 5e-6),
], ids=["qcl6", "marskar4", "marskar5", "v3", "v3_regions"])
def test_mfma_forms_vs_oracle(gpu, form, make, tol):
    """Rodrigues z-form with the generator product on MFMA (split FP16,
    FP32 accumulation) against the oracle: fields relative to their
    maximum, populations on the trace-1 scale."""
    spec = make()
    sol, recs = _run_hip(spec, form)
    ref = _oracle_records(sol)
    for r, a, b in zip(sol.get_results(), recs, ref):
        if r.get_name().startswith(("d", "inv")):
            err = float(np.max(np.abs(a - b)))
        else:
            err = rel_err(a, b)
        assert err <= tol, (r.get_name(), err)


@pytest.mark.parametrize("form", ["zform", "mfma", "mfma64", "mixed"])
def test_restructured_forms_from_noise(gpu, form):
    """tzenov2016 starts from the random 1e-15 V/m initial field.  At zero
    coherence the reference form's p = M_CP v^T (M d + d_eq) is exactly 0
    (v selects coherences that stay 0), while the rotated basis forms
    p = a . r + b from dense products, exact only to rounding: an absolute
    floor of a few 1e-10 V/m on E after 559 steps, above the 1e-15 V/m
    seed, far below the fields of a lasing device (1e5 V/m and more).  The
    populations follow the oracle to the forms' tolerances; E is held to
    that absolute floor."""
    spec = setups.tzenov2016(num_gridpoints=1024, endtime=20e-12)
    sol, recs = _run_hip(spec, form)
    ref = _oracle_records(sol)
    pop_tol = {"zform": 1e-12, "mfma64": MFMA_TOL, "mfma": MFMA_TOL,
               "mixed": 1e-6}[form]
    for r, a, b in zip(sol.get_results(), recs, ref):
        err = float(np.max(np.abs(a - b)))
        if r.get_name().startswith("d"):
            assert err <= pop_tol, (r.get_name(), err)
        else:
            assert err <= 2e-9, (r.get_name(), err)  # V/m


def test_mfma_batch_mixed_tiles(gpu):
    """Three six-level runs of uneven length in one context: tiles that
    straddle run boundaries take the VALU z-form kernel, the rest the MFMA
    kernel; every run matches its own oracle."""
    specs = [setups.qcl6(num_gridpoints=n, num_timesteps=1500, seed=s,
                         num_records=6, cw_ampl=3e6, name="q%d" % s)
             for s, n in ((1, 700), (2, 1100), (3, 333))]
    mb.material.clear_library()
    runs = [setups.build(s, clear_library=False) for s in specs]
    print("case mfma batch", flush=True)
    batch = mb.solver_hip_batch("hip-fdtd-6lvl-rodr-mfma", runs)
    info = batch.info()
    # MFMA tiles: the loaded span (owned cells plus the halo) lies inside
    # one run (runs follow each other with one pseudo cell in between)
    own, halo = info["owned_cells"], info["temporal_block"]
    spans, base = [], 0
    for sp in specs:
        spans.append((base, base + sp["num_gridpoints"]))
        base += sp["num_gridpoints"] + 1
    expect = sum(1 for t in range(info["num_tiles"])
                 if any(b0 <= t * own - halo and t * own - halo + 64 <= b1
                        for b0, b1 in spans))
    assert info["mfma_tiles"] == expect, (info, expect)
    batch.run()
    for r in range(len(specs)):
        ref = oracle.OracleRun(batch.problem(r)).run().records()
        for x, b in zip(batch.get_results(r), ref):
            a = x.get_data_real()
            err = (float(np.max(np.abs(a - b)))
                   if x.get_name().startswith("inv") else rel_err(a, b))
            assert err <= MFMA_TOL, (r, x.get_name(), err)


RED = SMALL[:3] + [
    ("noop", lambda: setups.maxwell_line(num_gridpoints=20000,
                                         num_timesteps=600)),
    ("cay5", lambda: dict(setups.marskar2011(5, num_gridpoints=400,
                                             endtime=0.08e-12),
                          algo="reg-cayley")),
]


@pytest.mark.parametrize("case", [c[0] for c in RED])
def test_red_order_matches_red_oracle(gpu, case):
    """hip-fdtd-red-* (density first, solver_cpu_fdtd_red) against the
    oracle's red loop (threads and overlap as the reference's defaults)."""
    spec = dict(RED)[case]()
    sol, recs = _run_hip(spec, order="red")
    assert sol.info()["step_order"] == capi.MBH_ORDER_RED
    ref = oracle.OracleRun(sol.problem()).run(order="red").records()
    for a, b in zip(recs, ref):
        assert a.shape == b.shape
        assert rel_err(a, b) <= FP64_TOL, (case, rel_err(a, b))


@pytest.mark.parametrize("tb,tile", [(1, 64), (7, 128), (24, 128)])
def test_red_order_independent_of_tiling(gpu, tb, tile):
    spec = setups.sit_2lvl(num_gridpoints=1500, endtime=0.5e-12,
                           num_records=16)
    _, base = _run_hip(spec, order="red", temporal_block=1, tile_cells=64)
    _, recs = _run_hip(spec, order="red", temporal_block=tb, tile_cells=tile)
    for a, b in zip(recs, base):
        assert np.array_equal(a, b)


def _v3_three_regions():
    """three-level medium between vacuum margins with point and full-grid
    population records: tiles at the region boundaries take the VALU body"""
    spec = setups.v3_rodr(num_gridpoints=1500, endtime=0.4e-12, num_records=9)
    spec["records"].append({"name": "d22", "type": 3, "row": 2, "col": 2,
                            "interval": spec["endtime"] / 8})
    spec["records"].append({"name": "d33", "type": 3, "row": 3, "col": 3,
                            "interval": 0.0, "position": 75e-6})
    return spec
