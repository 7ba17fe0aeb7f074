"""Benchmark of the MI355X Maxwell-Bloch hot path.

Workload (default, "sweep2"): 512 independent two-level self-induced
transparency runs (regularized Cayley; grid lengths seeded log-uniform in
2^12..2^17 cells, 19,039,104 cells in all; pulse areas U[0.5 pi, 4 pi]; T2
log-uniform in 10 fs..1 ns; dx, dt and the 49114 time steps shared), tile
geometry from the autotuner (mbsolve_amd.setups.sit_2lvl_sweep); candidate
forms: the closed Bloch form of the Cayley step with FP32 state storage
("mixed") and with FP64 storage ("zform").  One bench "step" is one pass:
every run advanced from its
initial state (mbh_reset) through all its time steps -- E update, sources,
H update and boundary, density update of every cell, record copy -- i.e. the
reference's solver run() (solver_cpu_fdtd.tpp:223-338) for all 512
scenarios.  --warmup W passes are untimed, then --steps K passes are timed
between device synchronisations.

Other workloads (--config): "sit2" (the single two-level SIT run, 8192
cells, 49114 steps), "v3" (three-level V system on Rodrigues, 65536 cells,
392939 steps), "qcl6" (six-level gain medium, 2^20 cells; a pass is the
first 20000 of its 4e6 steps unless --pass-steps says otherwise), "sweep6"
(64 six-level runs of 2^14..2^18 cells, passes as qcl6), "maxwell" (the pure
Maxwell stencil on 2^27 cells, 4000 steps).

Before the timed passes every candidate form and the comparison forms run
one pass each; their records are compared with those of a reference pass
(max|x - x_ref| / max|x_ref| per record and run): the plain FP64 form (the
reference's operation order) or, for the six-level workloads, the FP64
z-form (1e-13 from the oracle; the plain form is then timed on a short pass
for its speed).  The headline form is the fastest candidate within --tol.
A few runs are also compared with the CPU FP64 oracle (the restatement of
the reference's cpu-fdtd solver), where it finishes in seconds.  A bounded
sample of that oracle is timed on the host cores as the CPU baseline.

    python bench.py [--steps K] [--warmup W] [--config sweep2|sit2|v3|qcl6|
                    sweep6|maxwell] [--forms a,b] [--pass-steps N]
                    [--no-cpu] [--no-fp64] [--no-oracle]
"""
import argparse
import json
import os
import platform
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6
MFMA_PEAK_TFLOPS = 2500.0  # dense FP16 MFMA

# algorithmic bytes per cell update (state read + written once per step, no
# temporal reuse) by workload and storage precision; algorithmic flops
ALGO_BYTES = {("v3", "f32"): 80, ("v3", "f64"): 160,
              ("sit2", "f32"): 48, ("sit2", "f64"): 96,
              ("qcl6", "f32"): 296, ("qcl6", "f64"): 592,
              ("sweep2", "f32"): 48, ("sweep2", "f64"): 96,
              ("sweep6", "f32"): 296, ("sweep6", "f64"): 592,
              ("maxwell", "f32"): 16, ("maxwell", "f64"): 32}
ALGO_FLOPS = {"v3": 168, "sit2": 120, "qcl6": 2622, "sweep2": 120,
              "sweep6": 2622, "maxwell": 10}
# per workload: headline candidates, forms measured for comparison only,
# reference form for the errors ("" = plain FP64), default pass length
# (0 = the full run), steps of the plain FP64 timing when it is not the
# reference, runs checked against the CPU oracle ("smallest:k" or "all")
DEFAULTS = {
    "sweep2": ("mixed,zform", "", "", 0, 0, "smallest:4"),
    "sit2": ("mixed,zform", "", "", 0, 0, "all"),
    "v3": ("mixed,zform,mfma,mfma64", "", "", 0, 0, ""),
    "qcl6": ("mfma,mfma64", "zform", "zform", 20000, 400, ""),
    "sweep6": ("mfma,mfma64", "zform", "zform", 20000, 400, ""),
    "maxwell": ("mixed", "", "", 0, 0, ""),
}
STORAGE = {"mixed": "f32", "mfma": "f32", "zform": "f64", "mfma64": "f64",
           "": "f64"}
# forms whose arithmetic is FP64 throughout (the MFMA forms take their
# products in split FP16, "mixed" stores FP32)
EXACT64 = ("zform", "")
# v_mfma_f32_16x16x32_f16 per 16-cell block and step (split products x
# M-tiles x K chunks)
MFMA_PER_BLOCK = {6: 18, 5: 6, 4: 3, 3: 4}
MFMA_FLOPS = 2 * 16 * 16 * 32
LEVELS = {"qcl6": 6, "sweep6": 6, "v3": 3, "sit2": 2, "sweep2": 2,
          "maxwell": 0}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def make_specs(config, pass_steps):
    from mbsolve_amd import setups
    if config == "v3":
        return [setups.v3_rodr()], "3-level V system, Rodrigues, 65536 " \
            "cells, 150 um, 1.5 ps (392939 steps)"
    if config == "sit2":
        return [setups.sit_2lvl()], "2-level SIT, regularized Cayley, " \
            "8192 cells, 1.5 ps (49114 steps)"
    if config == "qcl6":
        return [setups.qcl6(num_timesteps=pass_steps)], \
            "6-level gain medium, Rodrigues, 2^20 cells, first %d of 4e6 " \
            "steps" % pass_steps
    if config == "sweep2":
        return setups.sit_2lvl_sweep(), "512 two-level SIT runs, " \
            "2^12..2^17 cells (19,039,104 in all), regularized Cayley, " \
            "49114 steps"
    if config == "maxwell":
        return [setups.maxwell_line()], "pure Maxwell FDTD (no density), " \
            "2^27 cells, 4000 steps"
    if config == "sweep6":
        return setups.qcl6_sweep(num_timesteps=pass_steps), "64 six-level " \
            "gain-medium runs, 2^14..2^18 cells, Rodrigues, first %d of 4e6 " \
            "steps" % pass_steps
    raise ValueError(config)


class Bench:
    """One solver (single run or batch) of one form, driven pass by pass."""

    def __init__(self, specs, form, opts):
        import mbsolve_amd.lib as mb
        from mbsolve_amd import setups
        self.form = form
        self.name = setups.solver_name(specs[0], form)
        if len(specs) == 1:
            dev, scen = setups.build(specs[0])
            self.sol = mb.create_hip_solver(self.name, dev, scen, **opts)
            self.batch = False
            self.n_t = int(self.sol.problem()["num_timesteps"])
        else:
            mb.material.clear_library()
            runs = [setups.build(s, clear_library=False) for s in specs]
            self.sol = mb.solver_hip_batch(self.name, runs, **opts)
            self.batch = True
            self.n_t = int(self.sol.problem(0)["num_timesteps"])
        self.info = self.sol.info()
        self.cells = int(self.info["total_cells"])

    def one_pass(self, steps, chunk_s=20.0):
        """Rewinds to the initial state and runs steps [0, steps); returns
        the device time (HIP events on the solver's stream around the
        launches) in ms after waiting for it.  A pass expected to take
        longer than about chunk_s runs in chunks of whole temporal blocks
        (the same launches), with a progress line after each."""
        self.sol.reset()
        tb = max(1, int(self.info["temporal_block"]))
        per_step = getattr(self, "_ms_per_step", 0.0)
        chunk = steps
        if per_step > 0.0 and per_step * steps > chunk_s * 1e3:
            chunk = max(tb, int(chunk_s * 1e3 / per_step) // tb * tb)
        dev, n = 0.0, 0
        while n < steps:
            m = min(steps, n + chunk)
            self.sol.run_steps(n, m)
            dev += self.sol.elapsed_ms()
            n = m
            if n < steps:
                log("    %d of %d steps, %.1f s device" % (n, steps,
                                                           dev * 1e-3))
        self._ms_per_step = dev / max(steps, 1)
        return dev

    def estimate(self, steps=None):
        """Device time per step from a short probe pass (for chunking)."""
        tb = max(1, int(self.info["temporal_block"]))
        k = min(self.n_t, max(8 * tb, 256) if steps is None else steps)
        ms = self.one_pass(k)
        self._ms_per_step = ms / k
        return self._ms_per_step

    def records(self):
        """[[record arrays of run r] for r] after the last pass."""
        self.sol.fetch_results()
        if self.batch:
            return [[x.get_data_real() for x in self.sol.get_results(r)]
                    for r in range(self.sol.num_runs())]
        return [[x.get_data_real() for x in self.sol.get_results()]]

    def record_names(self):
        res = self.sol.get_results(0) if self.batch else \
            self.sol.get_results()
        return [x.get_name() for x in res]

    def problem(self, r):
        return self.sol.problem(r) if self.batch else self.sol.problem()

    def close(self):
        self.sol = None


def rel_err(a, b):
    scale = float(np.max(np.abs(b))) if b.size else 0.0
    d = float(np.max(np.abs(a - b))) if b.size else 0.0
    return d / scale if scale > 0 else d


def error_table(recs, ref):
    """[run][record] relative errors."""
    return [[rel_err(a, b) for a, b in zip(ra, rb)]
            for ra, rb in zip(recs, ref)]


def summarize(errs, names):
    """Per record name: max over runs, median over runs, run of the max."""
    out = {}
    E = np.asarray(errs, dtype=float)
    for j, nm in enumerate(names):
        col = E[:, j]
        k = int(np.argmax(col))
        out[nm] = {"max": float(col[k]), "median": float(np.median(col)),
                   "worst_run": k}
    out["all_max"] = float(E.max()) if E.size else 0.0
    return out


def pick_headline(cands, results, tol, have_ref):
    """The fastest candidate form whose error is within tol (every
    candidate counts when none is) and whether one was (None without
    reference records, when no error is known)."""
    ok = [f for f in cands if "errs" not in results[f] or
          results[f]["errs"]["all_max"] <= tol]
    within_tol = bool(ok) if have_ref else None
    return max(ok or cands, key=lambda f: results[f]["value"]), within_tol


def fastest_exact64(results, fp64):
    """(form, value) of the fastest form with FP64 arithmetic throughout
    among the measured forms and the plain FP64 timing ("fp64"), or None."""
    exact64 = {f: results[f]["value"] for f in results if f in EXACT64}
    if fp64 is not None:
        exact64["fp64"] = fp64["value"]
    if not exact64:
        return None
    f = max(exact64, key=exact64.get)
    return f, exact64[f]


def timed_passes(b, steps, warmup, passes):
    """`warmup` untimed passes, then `passes` timed ones between device
    synchronisations (wall clock, and the summed device time of the
    launches)."""
    for i in range(warmup):
        ms = b.one_pass(steps)
        log("    warm-up pass %d/%d: %.1f ms device" % (i + 1, warmup, ms))
    dev = 0.0
    t0 = time.perf_counter()
    for i in range(passes):
        ms = b.one_pass(steps)  # waits for the pass's stop event
        dev += ms
        log("    timed pass %d/%d: %.1f ms device" % (i + 1, passes, ms))
    wall = time.perf_counter() - t0
    return wall, dev


def _cpu_info():
    model, phys = "unknown CPU", None
    try:
        out = subprocess.run(["lscpu"], stdout=subprocess.PIPE,
                             stderr=subprocess.DEVNULL, text=True).stdout
        kv = {}
        for line in out.splitlines():
            if ":" in line:
                k, v = line.split(":", 1)
                kv[k.strip()] = v.strip()
        model = kv.get("Model name", model)
        cps, sockets = kv.get("Core(s) per socket"), kv.get("Socket(s)")
        if cps and sockets and cps.isdigit() and sockets.isdigit():
            phys = int(cps) * int(sockets)
    except OSError:
        pass
    try:
        nproc = int(subprocess.run(["nproc"], stdout=subprocess.PIPE,
                                   text=True).stdout.strip())
    except (OSError, ValueError):
        nproc = os.cpu_count()
    return {"model": model, "nproc": nproc, "physical_cores": phys,
            "machine": platform.machine()}


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


def cpu_baseline(prob, seconds=15.0, reps=5):
    """Bounded sample of the CPU FP64 oracle on one run of the workload
    (its full grid): both CPU time loops of the reference (plain
    solver_cpu_fdtd and the redundant-halo solver_cpu_fdtd_red), each for
    `reps` timed pieces of as many steps as fit in about `seconds` / 2
    together; the faster loop's median rate is reported.  OpenMP threads:
    OMP_NUM_THREADS (the host cores this GPU's share of the machine holds),
    pinned to cores (OMP_PROC_BIND=close, OMP_PLACES=cores, set in main
    before the oracle library loads)."""
    info = _cpu_info()
    threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or \
        (info["physical_cores"] or os.cpu_count() or 1)
    nx = int(prob["num_gridpoints"])
    res = {o: _cpu_rate(prob, o, seconds / 2, reps, threads)
           for o in ("plain", "red")}
    best = max(res, key=lambda o: res[o][0])
    rate, per, total, warm = res[best]
    return {"value": rate, "unit": "cell-updates/s",
            "cores": threads, "kind": "port",
            "orders": {o: r[0] for o, r in res.items()},
            "host": info,
            "sample": "oracle (restated cpu-fdtd solver, FP64, built -O3 "
                      "-march=x86-64-v3; the faster of the plain (%.3g cu/s) "
                      "and red (%.3g cu/s) loops) on one %d-cell run, median "
                      "of %d pieces of %d steps after %d warm-up steps, %.1f "
                      "s, %d OpenMP threads pinned to cores (nproc %s, %s "
                      "physical cores, %s)"
                      % (res["plain"][0], res["red"][0], nx, reps, per, warm,
                         total, threads, info["nproc"],
                         info["physical_cores"], info["model"])}


def oracle_check(bench_recs, ref_recs, probs, which, threads):
    """Runs `which` through the CPU oracle; errors of the headline form and
    of the reference form against it, per record name (max over runs)."""
    import oracle
    out = {"runs": [], "cells": [], "form": None, "reference": None}
    ef, er = [], []
    for r in which:
        t0 = time.perf_counter()
        orc = oracle.OracleRun(probs[r]).run(threads=threads).records()
        log("    oracle run %d (%d cells): %.1f s" %
            (r, int(probs[r]["num_gridpoints"]),
             time.perf_counter() - t0))
        ef.append([rel_err(a, b) for a, b in zip(bench_recs[r], orc)])
        er.append([rel_err(a, b) for a, b in zip(ref_recs[r], orc)])
        out["runs"].append(r)
        out["cells"].append(int(probs[r]["num_gridpoints"]))
    out["form"] = ef
    out["reference"] = er
    return out


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


def load_profile(config, form, info):
    """Measured HBM traffic and VALU issue of the headline kernel at the
    tuned geometry, from profiles/traffic_<config>_<form>.json (rocprofv3
    FETCH_SIZE / WRITE_SIZE / SQ counters, see the file's command)."""
    path = os.path.join(ROOT, "profiles", "traffic_%s_%s.json"
                        % (config, form))
    if not os.path.exists(path):
        return None, None
    with open(path) as f:
        t = json.load(f)
    for g in t.get("geometries", [t]):
        if (g.get("temporal_block") == info["temporal_block"] and
                g.get("tile_cells") == info["tile_cells"]):
            return g, os.path.relpath(path, ROOT)
    return None, os.path.relpath(path, ROOT)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3,
                    help="timed passes (one pass = the whole workload "
                         "from its initial state through all its steps); "
                         "0 reports the headline form's selection pass")
    ap.add_argument("--warmup", type=int, default=1,
                    help="untimed passes before the timed ones")
    ap.add_argument("--config", default="sweep2", choices=sorted(DEFAULTS))
    ap.add_argument("--forms", default=None,
                    help="headline candidate forms (fastest within --tol)")
    ap.add_argument("--compare", default=None,
                    help="forms measured for comparison only")
    ap.add_argument("--ref", default=None,
                    help="reference form for the errors: '' (plain FP64) "
                         "or 'zform'")
    ap.add_argument("--pass-steps", type=int, default=None,
                    help="time steps per pass (default: the workload's)")
    ap.add_argument("--fp64-steps", type=int, default=None,
                    help="pass length of the plain FP64 timing when it is "
                         "not the reference")
    ap.add_argument("--tol", type=float, default=1e-5)
    ap.add_argument("--temporal-block", type=int, default=0)
    ap.add_argument("--tile-cells", type=int, default=0)
    ap.add_argument("--graph-launches", type=int, default=0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--no-oracle", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--per-run-out", default=None,
                    help="write the per-run error table (JSON) here")
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

    (d_cand, d_cmp, d_ref, d_pass, d_fp64_steps,
     d_oracle) = DEFAULTS[args.config]
    cands = [f for f in (args.forms if args.forms is not None
                         else d_cand).split(",") if f]
    compare = [f for f in (args.compare if args.compare is not None
                           else d_cmp).split(",") if f and f not in cands]
    ref_form = args.ref if args.ref is not None else d_ref
    pass_steps_req = args.pass_steps if args.pass_steps is not None \
        else d_pass
    fp64_steps = (args.fp64_steps if args.fp64_steps is not None
                  else d_fp64_steps)
    specs, workload = make_specs(args.config, pass_steps_req or 4000000)
    opts = dict(temporal_block=args.temporal_block,
                tile_cells=args.tile_cells,
                graph_launches=args.graph_launches)
    cpu_threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or \
        min(16, os.cpu_count() or 1)

    # reference pass: its records are what every form is compared with
    ref_recs, ref_names, fp64, probs = None, None, None, None
    if not args.no_fp64:
        log("reference form '%s' on %s" % (ref_form or "fp64", workload))
        b = Bench(specs, ref_form, opts)
        n_pass = min(pass_steps_req or b.n_t, b.n_t)
        if n_pass > 100000:
            b.estimate()
        ms = b.one_pass(n_pass)
        ref_recs, ref_names = b.records(), b.record_names()
        probs = [b.problem(r) for r in range(len(specs))]
        log("  %.1f ms device, %.4g cu/s, info %s"
            % (ms, b.cells * n_pass / (ms * 1e-3), b.info))
        if ref_form == "":
            fp64 = {"value": b.cells * n_pass / (ms * 1e-3), "steps": n_pass,
                    "info": b.info}
        b.close()
        if ref_form != "":
            log("plain FP64 form on a %d-step pass" % fp64_steps)
            b = Bench(specs, "", opts)
            ms = b.one_pass(fp64_steps)
            fp64 = {"value": b.cells * fp64_steps / (ms * 1e-3),
                    "steps": fp64_steps, "info": b.info}
            log("  %.1f ms device, %.4g cu/s" % (ms, fp64["value"]))
            b.close()

    # one pass of every form: speed and errors
    results, benches = {}, {}
    for form in cands + compare:
        log("form %s on %s" % (form, workload))
        b = Bench(specs, form, opts)
        n_pass = min(pass_steps_req or b.n_t, b.n_t)
        if n_pass > 100000:
            b.estimate()
        ms = b.one_pass(n_pass)
        r = {"value": b.cells * n_pass / (ms * 1e-3), "dev_ms": ms,
             "info": b.info, "steps": n_pass, "solver": b.name}
        if ref_recs is not None:
            errs = error_table(b.records(), ref_recs)
            r["errs"] = summarize(errs, ref_names)
            r["per_run"] = errs
        log("  %.1f ms device, %.4g cu/s, max rel err %s, info %s"
            % (ms, r["value"], r.get("errs", {}).get("all_max"), b.info))
        results[form] = r
        if form in cands:
            benches[form] = b
        else:
            b.close()

    head, within_tol = pick_headline(cands, results, args.tol,
                                     ref_recs is not None)
    for f in list(benches):
        if f != head:
            benches.pop(f).close()
    b = benches[head]
    n_pass = results[head]["steps"]
    cu_pass = b.cells * n_pass
    if args.steps > 0:
        log("headline form %s: %d warm-up + %d timed passes of %d steps"
            % (head, args.warmup, args.steps, n_pass))
        wall, dev_ms = timed_passes(b, n_pass, args.warmup, args.steps)
        value = cu_pass * args.steps / wall
    else:
        # --steps 0: the selection pass is the measurement (long passes,
        # e.g. a full 4e6-step horizon); device time of its launches
        dev_ms = results[head]["dev_ms"]
        wall = dev_ms * 1e-3
        value = results[head]["value"]
    passes = max(args.steps, 1)
    info = b.info
    sol_info = b.sol.info()
    launches = int(sol_info["launches"])
    # errors of the last timed pass (the headline's measured output)
    head_recs = b.records()
    errs = None
    if ref_recs is not None:
        per_run = error_table(head_recs, ref_recs)
        errs = summarize(per_run, ref_names)
        results[head]["per_run"] = per_run
    orc = None
    if ref_recs is not None and not args.no_oracle and d_oracle:
        if d_oracle == "all":
            which = list(range(len(specs)))
        else:
            k = int(d_oracle.split(":")[1])
            which = sorted(range(len(specs)),
                           key=lambda r: (int(probs[r]["num_gridpoints"]),
                                          r))[:k]
        log("CPU oracle on runs %s" % which)
        orc = oracle_check(head_recs, ref_recs, probs, which, cpu_threads)
        orc = {"runs": orc["runs"], "cells": orc["cells"],
               "form": summarize(orc["form"], ref_names),
               "reference": summarize(orc["reference"], ref_names)}
    prob0 = b.problem(0)
    b.close()

    storage = STORAGE[head]
    bpu = ALGO_BYTES[(args.config, storage)]
    dev_pass_ms = dev_ms / passes
    launch_ms = dev_pass_ms / max(launches, 1)
    units_per_launch = cu_pass / max(launches, 1)
    achieved = bpu * units_per_launch / (launch_ms * 1e-3) / 1e9
    lev = LEVELS[args.config]
    mfma_form = head.startswith("mfma")
    out = {
        "metric": "cell-updates/s, mixed vs FP64 GPU at <=1e-5 rel err; "
                  "% HBM/MFMA roofline",
        "value": value, "unit": "cell-updates/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": wall * 1e3 / passes, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None,
        "dtype": "f64 VALU + f16-split MFMA (f32 accumulate)" if mfma_form
        else "f64", "storage": storage, "data": "synthetic",
        "config": {"workload": workload, "runs": len(specs),
                   "cells": b.cells, "time_steps": n_pass,
                   "step": "one pass: every run from its initial state "
                           "through %d time steps" % n_pass,
                   "form": head, "solver": results[head]["solver"],
                   "temporal_block": info["temporal_block"],
                   "tile_cells": info["tile_cells"],
                   "owned_cells": info["owned_cells"],
                   "launches_per_pass": launches,
                   "graph_launches": info["graph_launches"]},
        "device_ms_per_step": dev_pass_ms,
        "within_tol": within_tol, "tol": args.tol,
        "errors_vs": None if ref_recs is None else
        ("plain FP64 GPU form" if ref_form == "" else
         "FP64 z-form (GPU)"),
        "errors": errs,
        "oracle_errors": orc,
        "roofline": {"bound": "hbm", "achieved": achieved,
                     "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "bytes_per_cell_update": bpu,
                     "units_per_launch": units_per_launch,
                     "launch_ms": launch_ms},
    }
    fpu = ALGO_FLOPS[args.config]
    out["compute"] = {"flops_per_cell_update": fpu,
                      "achieved_tflops": fpu * value / 1e12,
                      "fp64_peak_tflops": FP64_PEAK_TFLOPS,
                      "frac": fpu * value / 1e12 / FP64_PEAK_TFLOPS}
    prof, prof_path = load_profile(args.config, head, info)
    if prof is not None:
        out["roofline"]["traffic"] = prof.get("bytes_per_launch")
        bind = {"source": prof_path}
        if prof.get("bytes_per_launch"):
            steps_l = prof.get("steps_per_launch", info["temporal_block"])
            t_l = prof.get("kernel_ns", launch_ms * 1e6) * 1e-9
            bind["measured_hbm_gbs"] = prof["bytes_per_launch"] / t_l / 1e9
            bind["measured_hbm_frac"] = bind["measured_hbm_gbs"] / \
                HBM_PEAK_GBS
            bind["steps_per_launch"] = steps_l
        if prof.get("valu_issue_frac") is not None:
            bind["valu_issue_frac"] = prof["valu_issue_frac"]
            bind["bound"] = "valu_issue"
        if prof.get("mfma_busy_frac") is not None:
            bind["mfma_busy_frac"] = prof["mfma_busy_frac"]
        out["roofline"]["binding"] = bind
    if mfma_form:
        blocks = 4 * info["mfma_tiles"] * launches * info["temporal_block"]
        hw = blocks * MFMA_PER_BLOCK.get(lev, 0) * MFMA_FLOPS
        t = hw / (dev_pass_ms * 1e-3) / 1e12
        out["mfma"] = {"hardware_tflops": t, "peak_tflops": MFMA_PEAK_TFLOPS,
                       "frac": t / MFMA_PEAK_TFLOPS,
                       "instr": "v_mfma_f32_16x16x32_f16"}
    out["tool_rate"] = value * (n_pass - 1) / n_pass
    if fp64 is not None:
        out["fp64_value"] = fp64["value"]
        out["fp64_steps"] = fp64["steps"]
        out["speedup_vs_fp64"] = value / fp64["value"]
    ex64 = fastest_exact64(results, fp64)
    if ex64 is not None:
        out["speedup_vs_fastest_fp64_form"] = {
            "form": ex64[0], "value": ex64[1], "speedup": value / ex64[1]}
    out["forms"] = {f: {"value": r["value"], "storage": STORAGE[f],
                        "max_rel_err": (r["errs"]["all_max"]
                                        if "errs" in r else None),
                        "temporal_block": r["info"]["temporal_block"],
                        "tile_cells": r["info"]["tile_cells"],
                        "graph_launches": r["info"]["graph_launches"]}
                    for f, r in results.items()}
    if args.per_run_out:
        with open(args.per_run_out, "w") as f:
            json.dump({"records": ref_names,
                       "forms": {k: r.get("per_run")
                                 for k, r in results.items()}}, f)
    if not args.no_cpu:
        log("cpu baseline")
        out["cpu_baseline"] = cpu_baseline(prob0, seconds=args.cpu_seconds)
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
