"""Benchmark of the MI355X Maxwell-Bloch hot path.

Workload (default): the three-level V system on the Rodrigues propagator,
65536 cells over 150 um, run to 1.5 ps (392939 FDTD steps; mbsolve_amd.setups
.v3_rodr).  A "step" is one FDTD time step of the whole grid: E update,
sources, H update, density update of every cell, record copy.

Candidate fast forms (default: "mixed" = FP32 state storage with FP64
arithmetic and rotated-basis Rodrigues, "zform" = the same with FP64 storage)
and the plain FP64 form (reference operation order) are timed on the same
steps; the records of each fast form are compared with the FP64 form
(relative error max|x - x_fp64| / max|x_fp64|, per record) and the headline
value is the fastest form within the 1e-5 bound.  Device time comes from HIP events
on the solver's stream; the wall clock brackets the same region after a
device synchronisation.  A bounded sample of the CPU FP64 oracle (the
restatement of the reference's cpu-fdtd solver) is timed on the host cores.

    python bench.py [--steps K] [--warmup W] [--config v3|sit2|qcl6|sweep2]
                    [--no-cpu] [--no-fp64]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6

# algorithmic bytes per cell update (state read + write, no temporal reuse)
# and algorithmic flops per cell update, per workload and storage precision
ALGO_BYTES = {("v3", "f32"): 80, ("v3", "f64"): 160,
              ("sit2", "f32"): 48, ("sit2", "f64"): 96,
              ("qcl6", "f32"): 296, ("qcl6", "f64"): 592,
              ("sweep2", "f32"): 48, ("sweep2", "f64"): 96}
ALGO_FLOPS = {"v3": 168, "sit2": 120, "qcl6": 2622, "sweep2": 120}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def make_specs(config):
    from mbsolve_amd import setups
    if config == "v3":
        spec = setups.v3_rodr()
        return [spec], "configs[1]: 3-level V system, Rodrigues, 65536 " \
            "cells, 150 um, 1.5 ps"
    if config == "sit2":
        return [setups.sit_2lvl()], "configs[0]: 2-level SIT, " \
            "reg-Cayley, 8192 cells, 1.5 ps"
    if config == "qcl6":
        return [setups.qcl6(num_timesteps=20000)], "configs[2] shape: " \
            "6-level, 2^20 cells (20000 of 4e6 steps)"
    if config == "sweep2":
        return setups.sit_2lvl_sweep(), "configs[3]: 512-run 2-level sweep"
    raise ValueError(config)


def build_solver(specs, form, **opts):
    import mbsolve_amd.lib as mb
    from mbsolve_amd import setups
    if len(specs) == 1:
        dev, scen = setups.build(specs[0])
        name = setups.solver_name(specs[0], form)
        return mb.create_hip_solver(name, dev, scen, **opts)
    runs = []
    mb.material.clear_library()
    for s in specs:
        runs.append(setups.build(s, clear_library=False))
    name = setups.solver_name(specs[0], form)
    return mb.solver_hip_batch(name, runs, **opts)


def results_of(sol):
    if hasattr(sol, "num_runs"):
        out = []
        for r in range(sol.num_runs()):
            out.extend(x.get_data_real() for x in sol.get_results(r))
        return out
    return [x.get_data_real() for x in sol.get_results()]


def timed_run(sol, warmup, steps):
    """Warm-up steps [0, W), then the timed steps [W, W + K)."""
    if warmup:
        sol.run_steps(0, warmup)
        sol.fetch_results()  # waits for the device
    t0 = time.perf_counter()
    sol.run_steps(warmup, warmup + steps)
    dev_ms = sol.elapsed_ms()  # synchronises on the stop event
    wall = time.perf_counter() - t0
    sol.fetch_results()
    return dev_ms, wall


def cpu_baseline(specs, form_name, seconds=15.0):
    """Bounded sample of the CPU FP64 oracle on the same workload: the full
    grid of run 0 for as many steps as fit in about `seconds`."""
    import mbsolve_amd.lib as mb
    from mbsolve_amd import setups
    import oracle
    spec = specs[0]
    dev, scen = setups.build(spec)
    prob = mb.build_problem(setups.solver_name(spec), dev, scen)
    threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or \
        min(16, os.cpu_count() or 1)
    run = oracle.OracleRun(prob)
    nx = int(prob["num_gridpoints"])
    probe = 20
    t0 = time.perf_counter()
    run.run(0, probe, threads=threads)
    dt = time.perf_counter() - t0
    nsteps = int(max(probe, min(prob["num_timesteps"] - probe,
                                seconds / max(dt, 1e-9) * probe)))
    t0 = time.perf_counter()
    run.run(probe, probe + nsteps, threads=threads)
    el = time.perf_counter() - t0
    return {"value": nx * nsteps / el, "unit": "cell-updates/s",
            "cores": threads, "kind": "port",
            "sample": "oracle (restated cpu-fdtd plain solver, FP64) on the "
                      "full %d-cell grid, %d steps after %d warm-up steps, "
                      "%.1f s" % (nx, nsteps, probe, el)}


def measure(specs, form, warmup, steps, opts):
    """Builds the solver of `form`, runs warm-up + timed steps, returns
    (value cu/s, device ms, wall s, info, records)."""
    sol = build_solver(specs, form, **opts)
    prob0 = sol.problem() if not hasattr(sol, "num_runs") else sol.problem(0)
    n_t = int(prob0["num_timesteps"])
    warmup = min(warmup, n_t - 1)
    steps = min(steps if steps else n_t - warmup, n_t - warmup)
    dev_ms, wall = timed_run(sol, warmup, steps)
    info = sol.info()
    cells = int(info["total_cells"])
    recs = results_of(sol)
    del sol
    return {"value": cells * steps / (dev_ms * 1e-3), "dev_ms": dev_ms,
            "wall": wall, "info": info, "records": recs, "steps": steps,
            "warmup": warmup, "n_t": n_t, "cells": cells}


def rel_errs(recs, ref):
    errs = []
    for a, b in zip(recs, ref):
        scale = np.max(np.abs(b))
        d = np.max(np.abs(a - b))
        errs.append(float(d / scale) if scale > 0 else float(d))
    return errs


def copy_calibration(nbytes, reps=5):
    """Device-to-device hipMemcpy of `nbytes` (the runtime's plain copy
    kernel), for calibrating the FETCH_SIZE / WRITE_SIZE counters against a
    known byte count."""
    import ctypes as C
    hip = C.CDLL("libamdhip64.so")
    a, b = C.c_void_p(), C.c_void_p()
    if hip.hipMalloc(C.byref(a), C.c_size_t(nbytes)) or \
            hip.hipMalloc(C.byref(b), C.c_size_t(nbytes)):
        raise RuntimeError("hipMalloc failed")
    hip.hipMemset(a, 0, C.c_size_t(nbytes))
    hip.hipDeviceSynchronize()
    for _ in range(reps):
        t0 = time.perf_counter()
        hip.hipMemcpy(b, a, C.c_size_t(nbytes), 3)  # device to device
        hip.hipDeviceSynchronize()
        el = time.perf_counter() - t0
        log("copy %d bytes: %.3f ms, %.1f GB/s (read + write)"
            % (nbytes, el * 1e3, 2 * nbytes / el / 1e9))
    hip.hipFree(a)
    hip.hipFree(b)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=0,
                    help="timed steps (default: the rest of the run)")
    ap.add_argument("--warmup", type=int, default=2048)
    ap.add_argument("--config", default="v3")
    ap.add_argument("--forms", default="mixed,zform",
                    help="candidate fast forms; the headline is the fastest "
                         "one within the error bound of the FP64 form")
    ap.add_argument("--tol", type=float, default=1e-5)
    ap.add_argument("--temporal-block", type=int, default=0)
    ap.add_argument("--tile-cells", type=int, default=0)
    ap.add_argument("--graph-launches", type=int, default=0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--copy-calibration", type=int, default=0,
                    metavar="BYTES",
                    help="only run a device-to-device copy of BYTES")
    args = ap.parse_args()
    if args.copy_calibration:
        copy_calibration(args.copy_calibration)
        return

    if args.gpus != 1:
        print("bench.py: this build measures one GPU; runs on %d GPUs are "
              "not supported (each GPU would own a slab of the grid or a "
              "share of the batched runs)" % args.gpus, file=sys.stderr)
        sys.exit(2)

    import mbsolve_amd.lib as mb
    import mbsolve_amd.solverhip  # noqa: F401
    if mb.device_count() < 1:
        print("bench.py: no HIP device", file=sys.stderr)
        sys.exit(1)

    specs, workload = make_specs(args.config)
    opts = dict(temporal_block=args.temporal_block,
                tile_cells=args.tile_cells,
                graph_launches=args.graph_launches)

    forms = [f for f in args.forms.split(",") if f]
    results = {}
    for form in forms:
        log("measuring form", form, "on", workload)
        results[form] = measure(specs, form, args.warmup, args.steps, opts)
        r = results[form]
        log("  %s: %.3f ms device, %d launches, %.4g cu/s, info %s"
            % (form, r["dev_ms"], r["info"]["launches"], r["value"],
               r["info"]))
    ref = None
    if not args.no_fp64:
        log("measuring the FP64 reference-order form")
        ref = measure(specs, "", args.warmup, args.steps, opts)
        log("  fp64: %.3f ms device, %.4g cu/s" % (ref["dev_ms"],
                                                   ref["value"]))
        for form, r in results.items():
            r["errs"] = rel_errs(r["records"], ref["records"])
            log("  %s max rel err vs fp64: %s" % (form, r["errs"]))

    # headline: fastest form within the error bound (if measured)
    ok = [f for f in forms
          if ref is None or max(results[f]["errs"] or [0.0]) <= args.tol]
    pool = ok if ok else forms
    head = max(pool, key=lambda f: results[f]["value"])
    r = results[head]
    info = r["info"]
    storage = "f32" if head == "mixed" else "f64"
    launches = int(info["launches"])
    per_launch_units = r["cells"] * info["temporal_block"]
    launch_ms = r["dev_ms"] / max(launches, 1)
    bpu = ALGO_BYTES.get((args.config, storage))
    achieved = bpu * per_launch_units / (launch_ms * 1e-3) / 1e9
    fpu = ALGO_FLOPS.get(args.config)
    out = {
        "metric": "cell-updates/s, mixed vs FP64 GPU at <=1e-5 rel err; "
                  "% HBM/MFMA roofline",
        "value": r["value"], "unit": "cell-updates/s", "n_gpus": 1,
        "steps": r["steps"], "warmup": r["warmup"],
        "ms_per_step": r["dev_ms"] / r["steps"], "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "storage": storage, "data": "synthetic",
        "config": {"workload": workload, "cells": r["cells"],
                   "time_steps_total": r["n_t"], "form": head,
                   "temporal_block": info["temporal_block"],
                   "tile_cells": info["tile_cells"],
                   "owned_cells": info["owned_cells"],
                   "launches": launches},
        "wall_s": r["wall"],
        "roofline": {"bound": "hbm", "achieved": achieved,
                     "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "bytes_per_cell_update": bpu,
                     "units_per_launch": per_launch_units,
                     "launch_ms": launch_ms},
        "compute": {"flops_per_cell_update": fpu,
                    "achieved_tflops": fpu * r["value"] / 1e12,
                    "fp64_peak_tflops": FP64_PEAK_TFLOPS,
                    "frac": fpu * r["value"] / 1e12 / FP64_PEAK_TFLOPS},
        "forms": {f: {"value": results[f]["value"],
                      "max_rel_err_vs_fp64": (max(results[f]["errs"])
                                              if "errs" in results[f]
                                              else None),
                      "rel_err_per_record": results[f].get("errs"),
                      "temporal_block":
                          results[f]["info"]["temporal_block"],
                      "tile_cells": results[f]["info"]["tile_cells"]}
                  for f in forms},
    }
    traffic_file = os.path.join(ROOT, "profiles", "traffic_%s_%s.json"
                                % (args.config, head))
    if os.path.exists(traffic_file):
        with open(traffic_file) as f:
            t = json.load(f)
        if (t.get("temporal_block") == info["temporal_block"] and
                t.get("tile_cells") == info["tile_cells"]):
            out["roofline"]["traffic"] = t.get("bytes_per_launch")
    if ref is not None:
        out["fp64_value"] = ref["value"]
        out["speedup_vs_fp64"] = r["value"] / ref["value"]
        out["max_rel_err_vs_fp64"] = max(r["errs"]) if r["errs"] else 0.0
    if not args.no_cpu:
        log("cpu baseline")
        out["cpu_baseline"] = cpu_baseline(specs, head,
                                           seconds=args.cpu_seconds)
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
