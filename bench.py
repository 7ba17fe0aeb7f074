"""Benchmark of the MI355X Maxwell-Bloch hot path.

Workload (default, "v3"): the three-level V system on the Rodrigues
propagator, 65536 cells over 150 um, run to 1.5 ps (392939 FDTD steps;
mbsolve_amd.setups.v3_rodr).  A "step" is one FDTD time step of the whole
grid: E update, sources, H update, density update of every cell, record
copy.  Other workloads: "sit2" (two-level self-induced transparency,
regularized Cayley, 8192 cells, 1.5 ps), "qcl6" (six-level gain medium,
Rodrigues, 2^20 cells, the first 20000 of 4e6 steps), "sweep2" (512
two-level runs of 2^12..2^17 cells, 49114 shared steps), "sweep6" (64
six-level runs of 2^14..2^18 cells, the first 20000 of 4e6 steps),
"maxwell" (the pure Maxwell stencil on 2^27 cells, 4000 steps).

Candidate fast forms are timed on the same steps and their records compared
with an FP64 reference run (relative error max|x - x_ref| / max|x_ref|, per
record); the headline value is the fastest form within the 1e-5 bound.  The
reference is the plain FP64 form (reference operation order) or, for the
six-level workloads, the FP64 z-form (same arithmetic as the plain form up
to rounding, 1e-13 from the oracle); the plain FP64 form is then timed on
--fp64-steps steps for its speed.  Device time comes from HIP events on the
solver's stream; the wall clock brackets the same region after a device
synchronisation.  A bounded sample of the CPU FP64 oracle (the restatement
of the reference's cpu-fdtd solver) is timed on the host cores.

    python bench.py [--steps K] [--warmup W]
                    [--config v3|sit2|qcl6|sweep2|sweep6|maxwell]
                    [--forms a,b] [--no-cpu] [--no-fp64]
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
              ("sweep2", "f32"): 48, ("sweep2", "f64"): 96,
              ("sweep6", "f32"): 296, ("sweep6", "f64"): 592,
              ("maxwell", "f32"): 16, ("maxwell", "f64"): 32}
ALGO_FLOPS = {"v3": 168, "sit2": 120, "qcl6": 2622, "sweep2": 120,
              "sweep6": 2622, "maxwell": 10}
# per workload: default candidate forms, reference form for the errors,
# default steps of the plain FP64 timing (0: same as the candidates)
DEFAULTS = {"v3": ("mixed,zform,mfma,mfma64", "", 0),
            "sit2": ("mixed,zform", "", 0),
            "sweep2": ("mixed,zform", "", 0),
            "qcl6": ("mfma,mfma64,zform", "zform", 400),
            "sweep6": ("mfma,mfma64,zform", "zform", 400),
            "maxwell": ("mixed", "", 0)}
MFMA_PEAK_TFLOPS = 2500.0  # dense FP16 MFMA
# v_mfma_f32_16x16x32_f16 per 16-cell block and step: split products (3, or
# 4 with W - I in three pieces for three levels) x M-tiles x K chunks (six
# levels: 3 x 3 x 2)
MFMA_PER_BLOCK = {6: 18, 5: 6, 4: 3, 3: 4}
MFMA_FLOPS = 2 * 16 * 16 * 32


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def make_specs(config):
    from mbsolve_amd import setups
    if config == "v3":
        return [setups.v3_rodr()], "3-level V system, Rodrigues, 65536 " \
            "cells, 150 um, 1.5 ps (392939 steps)"
    if config == "sit2":
        return [setups.sit_2lvl()], "2-level SIT, regularized Cayley, " \
            "8192 cells, 1.5 ps (49114 steps)"
    if config == "qcl6":
        return [setups.qcl6(num_timesteps=20000)], "6-level gain medium, " \
            "Rodrigues, 2^20 cells, first 20000 of 4e6 steps"
    if config == "sweep2":
        return setups.sit_2lvl_sweep(), "512 two-level SIT runs, " \
            "2^12..2^17 cells, regularized Cayley, 49114 steps"
    if config == "maxwell":
        return [setups.maxwell_line()], "pure Maxwell FDTD (no density), " \
            "2^27 cells, 4000 steps"
    if config == "sweep6":
        return setups.qcl6_sweep(num_timesteps=20000), "64 six-level " \
            "gain-medium runs, 2^14..2^18 cells, Rodrigues, first 20000 " \
            "of 4e6 steps"
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


def timed_run(sol, warmup, steps, chunk_s=20.0):
    """Warm-up steps [0, W), then the timed steps [W, W + K) in chunks of
    about chunk_s seconds (one progress line each); returns the summed
    device time of the chunks and the wall time of the timed region."""
    if warmup:
        sol.run_steps(0, warmup)
        sol.fetch_results()  # waits for the device
        w_ms = sol.elapsed_ms()
        per_step = w_ms / warmup
    else:
        per_step = 0.0
    chunk = steps if per_step <= 0 else \
        max(1, min(steps, int(chunk_s * 1e3 / per_step)))
    dev_ms = 0.0
    t0 = time.perf_counter()
    done = 0
    while done < steps:
        n = min(chunk, steps - done)
        sol.run_steps(warmup + done, warmup + done + n)
        dev_ms += sol.elapsed_ms()  # synchronises on the stop event
        done += n
        if done < steps:
            log("    %d of %d steps, %.1f ms device" % (done, steps, dev_ms))
    wall = time.perf_counter() - t0
    sol.fetch_results()
    return dev_ms, wall


def _cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown CPU"


def _cpu_rate(prob, order, seconds, reps, threads):
    """Median cell-update rate of `reps` timed pieces of the oracle's
    `order` loop, as many steps as fit in about `seconds` together."""
    import oracle
    run = oracle.OracleRun(prob)
    nx = int(prob["num_gridpoints"])
    probe = 10
    run.run(0, probe, threads=threads, order=order)  # warm-up
    t0 = time.perf_counter()
    run.run(probe, 2 * probe, threads=threads, order=order)
    dt = time.perf_counter() - t0
    avail = int(prob["num_timesteps"]) - 2 * probe
    per = int(max(1, min(avail // reps,
                         seconds / reps / max(dt, 1e-9) * probe)))
    rates, total = [], 0.0
    n = 2 * probe
    for _ in range(reps):
        t0 = time.perf_counter()
        run.run(n, n + per, threads=threads, order=order)
        el = time.perf_counter() - t0
        rates.append(nx * per / el)
        total += el
        n += per
    rates.sort()
    return rates[len(rates) // 2], per, total, 2 * probe


def cpu_baseline(specs, form_name, seconds=15.0, reps=5):
    """Bounded sample of the CPU FP64 oracle on the same workload: the full
    grid of run 0, both CPU time loops of the reference (plain
    solver_cpu_fdtd and the redundant-halo solver_cpu_fdtd_red), each for
    `reps` timed pieces of as many steps as fit in about `seconds` / 2
    together; the faster loop's median rate is reported.  OpenMP threads are
    pinned to cores (OMP_PROC_BIND=close, OMP_PLACES=cores, set in main
    before the oracle library loads)."""
    import mbsolve_amd.lib as mb
    from mbsolve_amd import setups
    spec = specs[0]
    dev, scen = setups.build(spec)
    prob = mb.build_problem(setups.solver_name(spec), dev, scen)
    threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or \
        min(16, os.cpu_count() or 1)
    nx = int(prob["num_gridpoints"])
    res = {o: _cpu_rate(prob, o, seconds / 2, reps, threads)
           for o in ("plain", "red")}
    best = max(res, key=lambda o: res[o][0])
    rate, per, total, warm = res[best]
    return {"value": rate, "unit": "cell-updates/s",
            "cores": threads, "kind": "port",
            "orders": {o: r[0] for o, r in res.items()},
            "sample": "oracle (restated cpu-fdtd %s solver, FP64; the "
                      "faster of the plain and red loops, plain %.3g, red "
                      "%.3g cu/s) on the full %d-cell grid, median of %d "
                      "pieces of %d steps after %d warm-up steps, %.1f s, "
                      "%d OpenMP threads pinned to cores of %s"
                      % ("red" if best == "red" else "plain",
                         res["plain"][0], res["red"][0], nx, reps, per, warm,
                         total, threads, _cpu_model())}


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
    ap.add_argument("--config", default="v3", choices=sorted(DEFAULTS))
    ap.add_argument("--forms", default=None,
                    help="candidate fast forms; the headline is the fastest "
                         "one within the error bound of the FP64 reference")
    ap.add_argument("--ref", default=None,
                    help="reference form for the errors: '' (plain FP64) "
                         "or 'zform'")
    ap.add_argument("--fp64-steps", type=int, default=None,
                    help="steps of the plain FP64 timing when it is not "
                         "the reference (default per workload)")
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
    os.environ.setdefault("OMP_PROC_BIND", "close")
    os.environ.setdefault("OMP_PLACES", "cores")
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

    d_forms, d_ref, d_fp64_steps = DEFAULTS[args.config]
    forms = [f for f in (args.forms if args.forms is not None
                         else d_forms).split(",") if f]
    ref_form = args.ref if args.ref is not None else d_ref
    fp64_steps = (args.fp64_steps if args.fp64_steps is not None
                  else d_fp64_steps)
    specs, workload = make_specs(args.config)
    opts = dict(temporal_block=args.temporal_block,
                tile_cells=args.tile_cells,
                graph_launches=args.graph_launches)

    results = {}
    for form in forms:
        log("measuring form", form, "on", workload)
        results[form] = measure(specs, form, args.warmup, args.steps, opts)
        r = results[form]
        log("  %s: %.3f ms device, %d launches, %.4g cu/s, info %s"
            % (form, r["dev_ms"], r["info"]["launches"], r["value"],
               r["info"]))
    ref = None
    fp64 = None
    if not args.no_fp64:
        if ref_form in results:
            ref = results[ref_form]
        else:
            log("measuring the reference form '%s'" % ref_form)
            ref = measure(specs, ref_form, args.warmup, args.steps, opts)
            log("  ref: %.3f ms device, %.4g cu/s" % (ref["dev_ms"],
                                                      ref["value"]))
        for form, r in results.items():
            r["errs"] = rel_errs(r["records"], ref["records"])
            log("  %s max rel err vs the reference: %s"
                % (form, ["%.3g" % x for x in r["errs"]]))
        if ref_form == "":
            fp64 = ref
        else:
            log("timing the plain FP64 form on %s steps"
                % (fp64_steps or "the same"))
            fp64 = measure(specs, "", min(args.warmup, 64),
                           fp64_steps or args.steps, opts)
            log("  fp64: %.3f ms device, %.4g cu/s" % (fp64["dev_ms"],
                                                       fp64["value"]))

    # headline: fastest form within the error bound (if measured)
    ok = [f for f in forms
          if ref is None or max(results[f]["errs"] or [0.0]) <= args.tol]
    pool = ok if ok else forms
    head = max(pool, key=lambda f: results[f]["value"])
    r = results[head]
    info = r["info"]
    storage = "f32" if head in ("mixed", "mfma") else "f64"
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
                      "max_rel_err_vs_ref": (max(results[f]["errs"])
                                             if "errs" in results[f]
                                             else None),
                      "rel_err_per_record": results[f].get("errs"),
                      "temporal_block":
                          results[f]["info"]["temporal_block"],
                      "tile_cells": results[f]["info"]["tile_cells"]}
                  for f in forms},
        "reference_form": ("fp64" if ref_form == "" else ref_form)
        if ref is not None else None,
    }
    if head.startswith("mfma"):
        # matrix-core work of the tiles of one material with a density (the
        # tiles that straddle runs, regions or the grid ends take the VALU
        # z-form, vacuum tiles have no density)
        lev = {"qcl6": 6, "sweep6": 6, "v3": 3}.get(args.config, 0)
        blocks = 4 * info["mfma_tiles"] * launches * info["temporal_block"]
        hw = blocks * MFMA_PER_BLOCK.get(lev, 0) * MFMA_FLOPS
        t = hw / (r["dev_ms"] * 1e-3) / 1e12
        out["mfma"] = {"hardware_tflops": t, "peak_tflops": MFMA_PEAK_TFLOPS,
                       "frac": t / MFMA_PEAK_TFLOPS,
                       "mfma_per_tile_step": 4 * MFMA_PER_BLOCK.get(lev, 0),
                       "instr": "v_mfma_f32_16x16x32_f16, split FP16 "
                                "products (W - I in %d pieces)"
                                % (3 if lev <= 3 else 2)}
    traffic_file = os.path.join(ROOT, "profiles", "traffic_%s_%s.json"
                                % (args.config, head))
    if os.path.exists(traffic_file):
        with open(traffic_file) as f:
            t = json.load(f)
        # one measured geometry, or a list of them under "geometries"
        for g in t.get("geometries", [t]):
            if (g.get("temporal_block") == info["temporal_block"] and
                    g.get("tile_cells") == info["tile_cells"]):
                out["roofline"]["traffic"] = g.get("bytes_per_launch")
    # the figure mbsolve-tool prints (mbsolve-tool.cpp:693-695): N_x T / dt
    # per run time, i.e. n_t - 1 counted steps per n_t run steps, here for
    # a full run at the measured per-step rate
    out["tool_rate"] = r["value"] * (r["n_t"] - 1) / r["n_t"]
    if fp64 is not None:
        out["fp64_value"] = fp64["value"]
        out["fp64_steps"] = fp64["steps"]
        out["speedup_vs_fp64"] = r["value"] / fp64["value"]
    if ref is not None:
        out["max_rel_err_vs_ref"] = max(r["errs"]) if r["errs"] else 0.0
    if not args.no_cpu:
        log("cpu baseline")
        out["cpu_baseline"] = cpu_baseline(specs, head,
                                           seconds=args.cpu_seconds)
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
