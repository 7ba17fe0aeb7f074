"""CPU FP64 ORACLE -- test infrastructure only.

Restates the reference mbsolve's plain FDTD solver
(solver-cpu/include/mbsolve/solver-cpu/solver_cpu_fdtd.tpp) and its red
variant (solver_cpu_fdtd_red.tpp) so that the HIP path can be checked
against them:

* ``liboracle.so`` (fdtd_oracle.c): the two time loops and the per-cell
  density updates, run on an ``mbh_problem`` (the same struct the GPU gets);
* ``qm_setup``: an independent numpy/scipy restatement of the constructor's
  set-up (grid, coefficients, Cayley/Rodrigues constants, records, sources).

Only tests/, ``__graft_entry__.smoke()`` and the CPU-baseline leg of bench.py
use this package; the product never imports it.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

_lib = None


def load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise OSError("%s missing; run 'make -C oracle'" % LIB_PATH)
        lib = C.CDLL(LIB_PATH)
        dp = C.POINTER(C.c_double)
        lib.orc_run.restype = C.c_int
        lib.orc_run.argtypes = [C.c_void_p, dp, dp, dp, dp, dp, C.c_uint64,
                                C.c_uint64, C.c_int]
        lib.orc_run_red.restype = C.c_int
        lib.orc_run_red.argtypes = [C.c_void_p, dp, dp, dp, dp, dp,
                                    C.c_uint64, C.c_uint64, C.c_int, C.c_int]
        lib.orc_init_state.restype = None
        lib.orc_init_state.argtypes = [C.c_void_p, dp, dp, dp, dp]
        lib.orc_cayley_update.restype = None
        lib.orc_cayley_update.argtypes = [C.c_int, dp, dp, C.c_double, dp]
        lib.orc_rodr_update.restype = None
        lib.orc_rodr_update.argtypes = [C.c_int, dp, dp, C.c_double, dp]
        lib.orc_cv_population.restype = C.c_double
        lib.orc_cv_population.argtypes = [C.c_int, dp, C.c_int]
        _lib = lib
    return _lib


def _ptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


class OracleRun:
    """State of one oracle run on a problem dict (build_problem layout)."""

    def __init__(self, problem):
        from mbsolve_amd.capi import Problem  # struct layout of the ABI
        self.lib = load()
        self.problem = problem
        self.p = problem if isinstance(problem, Problem) else Problem(problem)
        d = self.p.dict
        nx = int(d["num_gridpoints"])
        self.nx = nx
        self.ds = len(d["d_init"]) // nx if nx else 0
        self.e = np.zeros(nx)
        self.h = np.zeros(nx + 1)
        self.pp = np.zeros(nx)
        self.d = np.zeros(max(nx * self.ds, 1))
        self.scratch = np.zeros(max(int(d["scratch_size"]), 1))
        self.lib.orc_init_state(C.byref(self.p.struct), _ptr(self.e),
                                _ptr(self.h), _ptr(self.pp), _ptr(self.d))

    def run(self, n_begin=0, n_end=None, threads=0, order="plain",
            overlap=0):
        """Steps [n_begin, n_end) in the plain order (solver_cpu_fdtd) or the
        red order (solver_cpu_fdtd_red; `overlap` = its OL, 0 = the
        reference's default)."""
        if n_end is None:
            n_end = int(self.p.dict["num_timesteps"])
        if order == "red":
            rc = self.lib.orc_run_red(
                C.byref(self.p.struct), _ptr(self.e), _ptr(self.h),
                _ptr(self.pp), _ptr(self.d), _ptr(self.scratch), n_begin,
                n_end, threads, overlap)
            if rc != 0:
                raise RuntimeError("orc_run_red failed (allocation or thread "
                                   "team size)")
        elif order == "plain":
            self.lib.orc_run(C.byref(self.p.struct), _ptr(self.e),
                             _ptr(self.h), _ptr(self.pp), _ptr(self.d),
                             _ptr(self.scratch), n_begin, n_end, threads)
        else:
            raise ValueError("order must be 'plain' or 'red'")
        return self

    def records(self):
        out = []
        for r in self.p.dict["records"]:
            o, n = int(r["offset_real"]), int(r["rows"]) * int(r["cols"])
            out.append(self.scratch[o:o + n].reshape(int(r["rows"]),
                                                     int(r["cols"])))
        return out

    def state(self):
        return {"e": self.e.copy(), "h": self.h.copy(), "p": self.pp.copy(),
                "d": self.d[:self.nx * self.ds].copy()}


def cayley_update(n, q, d, e):
    lib = load()
    q = np.ascontiguousarray(q, dtype=float)
    d = np.ascontiguousarray(d, dtype=float).copy()
    p = np.zeros(1)
    lib.orc_cayley_update(n, _ptr(q), _ptr(d), e, _ptr(p))
    return d, p[0]


def rodr_update(n, q, d, e):
    lib = load()
    q = np.ascontiguousarray(q, dtype=float)
    d = np.ascontiguousarray(d, dtype=float).copy()
    p = np.zeros(1)
    lib.orc_rodr_update(n, _ptr(q), _ptr(d), e, _ptr(p))
    return d, p[0]
