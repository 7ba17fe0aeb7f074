"""Failure handling of the solver factory, the host set-up and the C ABI
(all on the CPU; the C ABI validates problems before looking for a GPU)."""
import ctypes as C
import os
import re

import numpy as np
import pytest

import mbsolve_amd.lib as mb
import mbsolve_amd.solverhip  # noqa: F401
from mbsolve_amd import capi, setups

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_unknown_solver():
    spec = setups.sit_2lvl(num_gridpoints=64, endtime=1e-14)
    dev, scen = setups.build(spec)
    with pytest.raises(ValueError, match="Unknown solver cpu-fdtd-9lvl"):
        mb.solver.create_instance("cpu-fdtd-9lvl", dev, scen)
    with pytest.raises(ValueError, match="Unknown solver"):
        mb.build_problem("hip-fdtd-7lvl-rodr", dev, scen)


def test_names_registered_once():
    names = mb.solver.get_avail_solvers()
    for n in range(2, 7):
        assert "hip-fdtd-%dlvl-rodr" % n in names
        assert "hip-fdtd-%dlvl-rodr-mixed" % n in names
    assert "hip-fdtd-2lvl-reg-cayley" in names
    assert "hip-fdtd-2lvl-reg-cayley-zform" in names
    assert "hip-fdtd-noop" in names
    for n in range(3, 7):
        assert "hip-fdtd-%dlvl-rodr-mfma" % n in names
        assert "hip-fdtd-%dlvl-rodr-mfma64" % n in names
    assert "hip-fdtd-red-noop" in names
    for n in range(2, 7):
        assert "hip-fdtd-red-%dlvl-rodr" % n in names
        assert "hip-fdtd-red-%dlvl-reg-cayley" % n in names
    for bad in ("hip-fdtd-2lvl-rodr-mfma", "hip-fdtd-2lvl-reg-cayley-mfma",
                "hip-fdtd-3lvl-reg-cayley-zform", "hip-fdtd-noop-zform",
                "hip-fdtd-red-3lvl-rodr-mixed", "hip-fdtd-red-2lvl-rodr-zform",
                "hip-fdtd-red-noop-mixed"):
        assert bad not in names
    mb.load_hip_solvers()  # a second loader is harmless
    assert mb.solver.get_avail_solvers() == names


def test_level_mismatch():
    spec = setups.sit_2lvl(num_gridpoints=64, endtime=1e-14)
    dev, scen = setups.build(spec)
    with pytest.raises(ValueError, match="Number of energy levels"):
        mb.build_problem("hip-fdtd-3lvl-rodr", dev, scen)


def test_no_regions():
    dev = mb.device("empty")
    scen = mb.scenario("s", 16, 1e-15, mb.qm_operator([1.0, 0.0]))
    with pytest.raises(ValueError, match="No regions"):
        mb.build_problem("hip-fdtd-2lvl-reg-cayley", dev, scen)


def test_region_gap():
    spec = setups.sit_2lvl(num_gridpoints=64, endtime=1e-14)
    spec["regions"] = [spec["regions"][0], spec["regions"][2]]
    spec["regions"][1] = ("Vacuum right", "Vacuum", 100e-6, 150e-6)
    dev, scen = setups.build(spec)
    with pytest.raises(ValueError, match="region not found"):
        mb.build_problem("hip-fdtd-2lvl-reg-cayley", dev, scen)


def test_material_library():
    mb.material.clear_library()
    m = mb.material("X")
    mb.material.add_to_library(m)
    with pytest.raises(ValueError, match="already available"):
        mb.material.add_to_library(mb.material("X"))
    with pytest.raises(ValueError, match="not available"):
        mb.material.get_from_library("Y")


def test_unknown_record_type():
    with pytest.raises(ValueError, match="Unknown record type"):
        mb.record("zz", 0.0)


def test_singular_time_independent_matrix():
    """Rodrigues needs M d_in = d_eq to be solvable
    (algo_lindblad_cvr_rodr.hpp:199-205): with all dynamics switched off
    but a non-zero equilibrium vector the set-up must refuse."""
    q = {"num_levels": 2, "carrier_density": 1e24, "h_diag": [0.0, 0.0],
         "h_off": [0.0], "mu_diag": [0.0, 0.0], "mu_off": [0.0],
         "scattering": [[0.0, 0.0], [1e9, 0.0]], "pure_dephasing": [0.0]}
    spec = {"name": "sing", "algo": "cvr-rodr", "num_levels": 2,
            "materials": {"M": {"qm": q}},
            "regions": [("r", "M", 0.0, 1e-6)], "num_gridpoints": 8,
            "endtime": 1e-15, "rho_init": ([1.0, 0.0], [])}
    dev, scen = setups.build(spec)
    # M is singular here only if relaxation has a null direction matching
    # d_eq; the solver accepts any solvable system
    try:
        mb.build_problem("hip-fdtd-2lvl-rodr", dev, scen)
    except ValueError as e:
        assert "invertible" in str(e)


def test_source_outside_device():
    spec = setups.sit_2lvl(num_gridpoints=64, endtime=1e-14)
    spec["sources"][0]["position"] = 1.0
    dev, scen = setups.build(spec)
    with pytest.raises(ValueError, match="Source position"):
        mb.build_problem("hip-fdtd-2lvl-reg-cayley", dev, scen)


# ---- C ABI ----------------------------------------------------------------
def _header_functions():
    text = open(os.path.join(ROOT, "include", "mbsolve_hip.h")).read()
    return sorted(set(re.findall(r"\b(mbh_[a-z_]+)\s*\(", text)))


def test_capi_exports_every_declared_symbol():
    lib = C.CDLL(capi.LIB_PATH)
    names = _header_functions()
    assert set(names) == set(capi.EXPORTED)
    for n in names:
        assert hasattr(lib, n), n
    capi.load()
    assert capi.load().mbh_version().decode().startswith("mbsolve-hip")


def _good_problem():
    spec = setups.sit_2lvl(num_gridpoints=64, endtime=2e-14, num_records=4)
    dev, scen = setups.build(spec)
    return mb.build_problem("hip-fdtd-2lvl-reg-cayley", dev, scen)


@pytest.mark.parametrize("mutate,msg", [
    (lambda d: d.__setitem__("qm_stride", 12), "qm_stride"),
    (lambda d: d.__setitem__("num_levels", 9), "num_levels"),
    (lambda d: d.__setitem__("algo", 7), "unknown algo"),
    (lambda d: d["mat_idx"].__setitem__(3, 5), "mat_idx"),
    (lambda d: d["source_x"].__setitem__(0, 10 ** 6), "outside the grid"),
    (lambda d: d["records"][0].__setitem__("cols", 10 ** 6), "inconsistent"),
])
def test_capi_rejects_inconsistent_problems(mutate, msg):
    d = _good_problem()
    d["mat_idx"] = np.array(d["mat_idx"])
    d["source_x"] = np.array(d["source_x"])
    mutate(d)
    with pytest.raises(capi.MBHError, match=msg) as ei:
        capi.Context([d])
    assert ei.value.code == 1  # MBH_ERR_INVALID


def test_capi_null_arguments():
    lib = capi.load()
    out = C.c_void_p()
    assert lib.mbh_create(None, 0, None, C.byref(out)) == 1
    assert b"no runs" in lib.mbh_last_error(None)
    assert lib.mbh_run(None, 0, 1) == 1
    lib.mbh_destroy(None)


@pytest.mark.parametrize("form,code,msg", [
    ("mfma", 4, "MFMA form"),     # Cayley has no MFMA form
    (9, 1, "unknown form"),
])
def test_capi_rejects_unsupported_forms(form, code, msg):
    d = _good_problem()
    with pytest.raises(capi.MBHError, match=msg) as ei:
        capi.Context([d], capi.options(form))
    assert ei.value.code == code


@pytest.mark.parametrize("form,order,code,msg", [
    ("mixed", "red", 4, "red step order"),   # red order: FP64 form only
    ("zform64", "red", 4, "red step order"),
    ("fp64", 2, 1, "unknown step order"),
])
def test_capi_rejects_unsupported_orders(form, order, code, msg):
    d = _good_problem()
    with pytest.raises(capi.MBHError, match=msg) as ei:
        capi.Context([d], capi.options(form, order=order))
    assert ei.value.code == code


def test_red_names_parse_without_gpu():
    spec = setups.v3_rodr(num_gridpoints=64, endtime=2e-14, num_records=4)
    dev, scen = setups.build(spec)
    red = mb.build_problem(setups.solver_name(spec, order="red"), dev, scen)
    plain = mb.build_problem(setups.solver_name(spec), dev, scen)
    assert setups.solver_name(spec, order="red") == "hip-fdtd-red-3lvl-rodr"
    assert np.array_equal(red["qm_consts"], plain["qm_consts"])
    with pytest.raises(ValueError, match="Unknown solver"):
        mb.build_problem("hip-fdtd-red-3lvl-rodr-mixed", dev, scen)


def test_capi_mfma_needs_three_levels():
    """A two-level Rodrigues problem is refused by the MFMA form."""
    dev, scen = setups.build(setups.sit_2lvl(num_gridpoints=64,
                                             endtime=2e-14))
    d = mb.build_problem("hip-fdtd-2lvl-rodr-zform", dev, scen)
    with pytest.raises(capi.MBHError, match="MFMA form") as ei:
        capi.Context([d], capi.options("mfma64"))
    assert ei.value.code == 4
