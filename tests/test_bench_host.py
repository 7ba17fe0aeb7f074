"""Host logic of bench.py on the CPU: the error measure, the choice of the
headline form, the comparison with the fastest FP64 form, the profile
lookup behind the roofline, and the committed bench lines against the
output contract (keys, derived figures, their profile sources)."""
import glob
import json
import os
import subprocess
import sys

import numpy as np
import pytest

import bench

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_rel_err_is_max_abs_difference_over_max_abs_reference():
    b = np.array([1.0, -4.0, 2.0])
    a = np.array([1.5, -4.0, 1.0])
    assert bench.rel_err(a, b) == pytest.approx(1.0 / 4.0)
    # an all-zero reference: the absolute difference
    assert bench.rel_err(np.array([0.0, 3e-9]), np.zeros(2)) == 3e-9
    assert bench.rel_err(np.zeros(0), np.zeros(0)) == 0.0


def test_summarize_max_median_and_worst_run():
    errs = [[1e-6, 4e-7], [3e-6, 1e-7], [2e-6, 9e-7]]
    s = bench.summarize(errs, ["inv12", "e"])
    assert s["inv12"] == {"max": 3e-6, "median": 2e-6, "worst_run": 1}
    assert s["e"] == {"max": 9e-7, "median": 4e-7, "worst_run": 2}
    assert s["all_max"] == 3e-6


def _res(value, err=None):
    r = {"value": value}
    if err is not None:
        r["errs"] = {"all_max": err}
    return r


def test_headline_is_the_fastest_form_within_the_tolerance():
    results = {"mixed": _res(5.7e11, 1.1e-5), "zform": _res(5.6e11, 5e-11)}
    head, within = bench.pick_headline(["mixed", "zform"], results, 1e-5,
                                       True)
    assert (head, within) == ("zform", True)
    results["mixed"] = _res(5.7e11, 7e-6)
    assert bench.pick_headline(["mixed", "zform"], results, 1e-5,
                               True) == ("mixed", True)


def test_headline_without_a_form_within_the_tolerance_says_so():
    results = {"mfma": _res(1.7e10, 2e-5), "mfma64": _res(1.9e10, 3e-5)}
    assert bench.pick_headline(["mfma", "mfma64"], results, 1e-5,
                               True) == ("mfma64", False)
    # no reference pass: no error known, the fastest form, no verdict
    results = {"mixed": _res(4.3e12)}
    assert bench.pick_headline(["mixed"], results, 1e-5,
                               False) == ("mixed", None)


def test_fastest_fp64_form_counts_only_fp64_arithmetic():
    """The MFMA forms take their products in split FP16 (mfma64 stores FP64
    but is not FP64 arithmetic) and `mixed` stores FP32: neither counts."""
    results = {"mfma": _res(1.6e10), "mfma64": _res(1.9e10),
               "zform": _res(1.3e10)}
    assert bench.fastest_exact64(results, {"value": 1.6e7}) == \
        ("zform", 1.3e10)
    # the plain FP64 timing when no FP64 form was among the measured ones
    assert bench.fastest_exact64({"mixed": _res(4.3e12)},
                                 {"value": 2.2e12}) == ("fp64", 2.2e12)
    assert bench.fastest_exact64({"mixed": _res(4.3e12)}, None) is None


def test_profile_lookup_matches_the_tuned_geometry():
    info = {"temporal_block": 24, "tile_cells": 256}
    g, src = bench.load_profile("sweep2", "mixed", info)
    assert src == os.path.join("profiles", "traffic_sweep2_mixed.json")
    assert g["temporal_block"] == 24 and g["tile_cells"] == 256
    g, src = bench.load_profile("sweep2", "mixed",
                                {"temporal_block": 5, "tile_cells": 256})
    assert g is None and src is not None
    assert bench.load_profile("nosuch", "mixed", info) == (None, None)


TRAFFIC = sorted(glob.glob(os.path.join(ROOT, "profiles", "traffic_*.json")))


@pytest.mark.parametrize("path", TRAFFIC,
                         ids=[os.path.basename(p) for p in TRAFFIC])
def test_traffic_profiles_are_consistent(path):
    """bytes per launch = FETCH_SIZE x the copy-calibrated factor + WRITE_SIZE
    (KiB medians), as the roofline's measured traffic is computed; the VALU
    issue share = SQ_INSTS_VALU x 4 cycles / (1024 SIMDs x 2.4 GHz x kernel
    time).  (The share can exceed 1 where 32-bit instructions issue in
    fewer than 4 cycles: the pure Maxwell kernel at 4 cells per lane.)"""
    t = json.load(open(path))
    fs, ws = t["fetch_scale"], t["write_scale"]
    for g in t["geometries"]:
        want = (g["fetch_kib_median"] * fs + g["write_kib_median"] * ws) * 1024
        assert g["bytes_per_launch"] == pytest.approx(want, rel=1e-6), g
        issue = g["sq_insts_valu_median"] * 4 / (
            1024 * 2.4e9 * g["kernel_ns"] * 1e-9)
        assert g["valu_issue_frac"] == pytest.approx(issue, abs=6e-5), g


BENCH = sorted(os.path.join(ROOT, "profiles", "bench_%s.json" % c)
               for c in bench.DEFAULTS)
KEYS = ("metric", "value", "unit", "n_gpus", "steps", "warmup",
        "ms_per_step", "higher_is_better", "scaling", "vs_baseline", "dtype",
        "data", "config", "roofline", "cpu_baseline")


@pytest.mark.parametrize("path", BENCH,
                         ids=[os.path.basename(p) for p in BENCH])
def test_committed_bench_lines_keep_the_contract(path):
    d = json.load(open(path))
    for k in KEYS:
        assert k in d, k
    assert d["n_gpus"] == 1 and d["higher_is_better"] is True
    assert "workload" in d["config"]
    r = d["roofline"]
    assert r["frac"] == pytest.approx(r["achieved"] / r["peak"])
    assert r["achieved"] == pytest.approx(
        r["bytes_per_cell_update"] * r["units_per_launch"] /
        (r["launch_ms"] * 1e-3) / 1e9)
    b = r.get("binding")
    if b is not None:
        assert os.path.exists(os.path.join(ROOT, b["source"]))
    c = d["cpu_baseline"]
    assert c["kind"] in ("port", "reference") and c["cores"] >= 1
    assert c["value"] > 0 and c["sample"]
    if d.get("within_tol") is not None and d.get("errors"):
        assert d["within_tol"] == (d["errors"]["all_max"] <= d["tol"])
    fx = d.get("speedup_vs_fastest_fp64_form")
    if fx is not None:
        assert fx["form"] in ("zform", "fp64")
        assert fx["speedup"] == pytest.approx(d["value"] / fx["value"])


def test_more_than_one_gpu_exits_with_a_message():
    p = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"),
                        "--gpus", "2"], capture_output=True, text=True,
                       timeout=120)
    assert p.returncode == 2
    assert "one GPU" in p.stderr and p.stdout == ""
