"""ctypes binding of the C ABI declared in include/mbsolve_hip.h.

This is what a Python host (or any FFI) binds: plain structs and pointers,
integer status codes, mbh_last_error() for the message.  ``problem_from_dict``
turns the host set-up produced by ``mbsolve_amd.lib.build_problem`` (or the
``problem()`` of a hip solver) into an ``mbh_problem``.
"""
import ctypes as C
import os

import numpy as np

from . import PKG_DIR

LIB_PATH = os.path.join(PKG_DIR, "libmbsolve_hip.so")

MBH_ABI_VERSION = 1
MBH_OK = 0
MBH_ALGO_NOOP, MBH_ALGO_REG_CAYLEY, MBH_ALGO_CVR_RODR = 0, 1, 2
MBH_FORM_FP64, MBH_FORM_MIXED, MBH_FORM_ZFORM64 = 0, 1, 2
MBH_FORM_MFMA, MBH_FORM_MFMA64 = 3, 4
MBH_ORDER_PLAIN, MBH_ORDER_RED = 0, 1
ORDERS = {"plain": MBH_ORDER_PLAIN, "red": MBH_ORDER_RED}
FORMS = {"fp64": MBH_FORM_FP64, "mixed": MBH_FORM_MIXED,
         "zform64": MBH_FORM_ZFORM64, "mfma": MBH_FORM_MFMA,
         "mfma64": MBH_FORM_MFMA64}

EXPORTED = ["mbh_version", "mbh_device_count", "mbh_default_options",
            "mbh_create", "mbh_run", "mbh_reset", "mbh_synchronize",
            "mbh_elapsed_ms",
            "mbh_fetch_records", "mbh_fetch_state",
            "mbh_fetch_state_noalloc", "mbh_get_info",
            "mbh_destroy", "mbh_last_error"]


class mbh_source(C.Structure):
    _fields_ = [("type", C.c_uint32), ("x_idx", C.c_uint32)]


class mbh_record(C.Structure):
    _fields_ = [("type", C.c_uint32), ("row_idx", C.c_uint32),
                ("col_idx", C.c_uint32), ("is_complex", C.c_uint32),
                ("rows", C.c_uint64), ("cols", C.c_uint64),
                ("position_idx", C.c_uint64), ("interval_ratio", C.c_double),
                ("offset_real", C.c_uint64), ("offset_imag", C.c_uint64)]


_dp = C.POINTER(C.c_double)


class mbh_problem(C.Structure):
    _fields_ = [("abi_version", C.c_uint32), ("algo", C.c_uint32),
                ("num_levels", C.c_uint32), ("num_materials", C.c_uint32),
                ("num_gridpoints", C.c_uint64),
                ("num_timesteps", C.c_uint64), ("dx_inv", C.c_double),
                ("fdtd_consts", _dp), ("qm_consts", _dp),
                ("qm_stride", C.c_uint64), ("zform_consts", _dp),
                ("zform_stride", C.c_uint64),
                ("mat_idx", C.POINTER(C.c_uint32)), ("e_init", _dp),
                ("h_init", _dp), ("d_init", _dp),
                ("num_sources", C.c_uint32),
                ("sources", C.POINTER(mbh_source)), ("source_data", _dp),
                ("num_records", C.c_uint32),
                ("records", C.POINTER(mbh_record)),
                ("scratch_size", C.c_uint64)]


class mbh_options(C.Structure):
    _fields_ = [("form", C.c_uint32), ("temporal_block", C.c_uint32),
                ("tile_cells", C.c_uint32), ("device", C.c_int32),
                ("graph_launches", C.c_uint32), ("step_order", C.c_uint32),
                ("flags", C.c_uint32), ("reserved", C.c_uint32 * 1)]

MBH_OPT_GENERIC = 1


class mbh_info(C.Structure):
    _fields_ = [("form", C.c_uint32), ("temporal_block", C.c_uint32),
                ("tile_cells", C.c_uint32), ("owned_cells", C.c_uint32),
                ("num_tiles", C.c_uint64), ("total_cells", C.c_uint64),
                ("launches", C.c_uint64), ("state_bytes", C.c_uint64),
                ("graph_launches", C.c_uint32), ("step_order", C.c_uint32),
                ("mfma_tiles", C.c_uint32), ("flags", C.c_uint32)]

MBH_INFO_HZ_ZERO = 1
MBH_INFO_BLOCH_REAL = 2
MBH_NOALLOC_SCRATCH_BYTES = 64 << 20


_lib = None


def load():
    """Loads libmbsolve_hip.so and declares the prototypes."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise OSError("%s missing; run 'python -m mbsolve_amd.build'"
                      % LIB_PATH)
    lib = C.CDLL(LIB_PATH)
    vp = C.c_void_p
    lib.mbh_version.restype = C.c_char_p
    lib.mbh_version.argtypes = []
    lib.mbh_device_count.restype = C.c_int
    lib.mbh_device_count.argtypes = [C.POINTER(C.c_int)]
    lib.mbh_default_options.restype = None
    lib.mbh_default_options.argtypes = [C.POINTER(mbh_options)]
    lib.mbh_create.restype = C.c_int
    lib.mbh_create.argtypes = [C.POINTER(mbh_problem), C.c_uint32,
                               C.POINTER(mbh_options), C.POINTER(vp)]
    lib.mbh_run.restype = C.c_int
    lib.mbh_run.argtypes = [vp, C.c_uint64, C.c_uint64]
    lib.mbh_reset.restype = C.c_int
    lib.mbh_reset.argtypes = [vp]
    lib.mbh_synchronize.restype = C.c_int
    lib.mbh_synchronize.argtypes = [vp]
    lib.mbh_elapsed_ms.restype = C.c_int
    lib.mbh_elapsed_ms.argtypes = [vp, C.POINTER(C.c_double)]
    lib.mbh_fetch_records.restype = C.c_int
    lib.mbh_fetch_records.argtypes = [vp, C.c_uint32, _dp, C.c_uint64]
    lib.mbh_fetch_state.restype = C.c_int
    lib.mbh_fetch_state.argtypes = [vp, C.c_uint32, _dp, _dp, _dp, _dp]
    lib.mbh_fetch_state_noalloc.restype = C.c_int
    lib.mbh_fetch_state_noalloc.argtypes = [vp, C.c_uint32, _dp, _dp, _dp,
                                            _dp]
    lib.mbh_get_info.restype = C.c_int
    lib.mbh_get_info.argtypes = [vp, C.POINTER(mbh_info)]
    lib.mbh_destroy.restype = None
    lib.mbh_destroy.argtypes = [vp]
    lib.mbh_last_error.restype = C.c_char_p
    lib.mbh_last_error.argtypes = [vp]
    _lib = lib
    return lib


class MBHError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("mbsolve_hip error %d: %s" % (code, msg))
        self.code = code


def _f64(a):
    return np.ascontiguousarray(np.asarray(a, dtype=np.float64))


class Problem:
    """Owns the numpy buffers an mbh_problem points to."""

    def __init__(self, d):
        self.keep = []
        p = mbh_problem()
        p.abi_version = MBH_ABI_VERSION
        p.algo = int(d["algo"])
        p.num_levels = int(d["num_levels"])
        p.num_materials = int(d["num_materials"])
        p.num_gridpoints = int(d["num_gridpoints"])
        p.num_timesteps = int(d["num_timesteps"])
        p.dx_inv = float(d["dx_inv"])

        def dptr(a):
            if a is None or len(a) == 0:
                return None
            a = _f64(a)
            self.keep.append(a)
            return a.ctypes.data_as(_dp)

        p.fdtd_consts = dptr(d["fdtd_consts"])
        p.qm_consts = dptr(d["qm_consts"])
        p.qm_stride = int(d["qm_stride"])
        p.zform_consts = dptr(d.get("zform_consts"))
        p.zform_stride = int(d.get("zform_stride", 0))
        mi = np.ascontiguousarray(np.asarray(d["mat_idx"], dtype=np.uint32))
        self.keep.append(mi)
        p.mat_idx = mi.ctypes.data_as(C.POINTER(C.c_uint32))
        p.e_init = dptr(d["e_init"])
        p.h_init = dptr(d["h_init"])
        p.d_init = dptr(d["d_init"])
        ns = len(d["source_type"])
        src = (mbh_source * max(ns, 1))()
        for k in range(ns):
            src[k].type = int(d["source_type"][k])
            src[k].x_idx = int(d["source_x"][k])
        self.keep.append(src)
        p.num_sources = ns
        p.sources = C.cast(src, C.POINTER(mbh_source)) if ns else None
        p.source_data = dptr(d["source_data"])
        recs = d["records"]
        nr = len(recs)
        rr = (mbh_record * max(nr, 1))()
        for k, r in enumerate(recs):
            for f, _ in mbh_record._fields_:
                setattr(rr[k], f, r[f])
        self.keep.append(rr)
        p.num_records = nr
        p.records = C.cast(rr, C.POINTER(mbh_record)) if nr else None
        p.scratch_size = int(d["scratch_size"])
        self.struct = p
        self.dict = d


def problem_from_dict(d):
    return Problem(d)


def options(form="fp64", temporal_block=0, tile_cells=0, graph_launches=0,
            device=-1, order="plain", generic=False):
    lib = load()
    o = mbh_options()
    lib.mbh_default_options(C.byref(o))
    o.form = FORMS[form] if isinstance(form, str) else int(form)
    o.temporal_block = temporal_block
    o.tile_cells = tile_cells
    o.graph_launches = graph_launches
    o.device = device
    o.step_order = ORDERS[order] if isinstance(order, str) else int(order)
    o.flags = MBH_OPT_GENERIC if generic else 0
    return o


class Context:
    """One mbh_ctx: several runs advanced together on one GPU."""

    def __init__(self, problems, opts=None):
        lib = load()
        self.lib = lib
        self.problems = [p if isinstance(p, Problem) else Problem(p)
                         for p in problems]
        arr = (mbh_problem * len(self.problems))(
            *[p.struct for p in self.problems])
        self.ctx = C.c_void_p()
        o = opts if opts is not None else options()
        rc = lib.mbh_create(arr, len(self.problems), C.byref(o),
                            C.byref(self.ctx))
        if rc != MBH_OK:
            raise MBHError(rc, lib.mbh_last_error(None).decode())

    def _check(self, rc):
        if rc != MBH_OK:
            raise MBHError(rc, self.lib.mbh_last_error(self.ctx).decode())

    def run(self, n_begin, n_end):
        self._check(self.lib.mbh_run(self.ctx, n_begin, n_end))

    def reset(self):
        self._check(self.lib.mbh_reset(self.ctx))

    def synchronize(self):
        self._check(self.lib.mbh_synchronize(self.ctx))

    def elapsed_ms(self):
        ms = C.c_double()
        self._check(self.lib.mbh_elapsed_ms(self.ctx, C.byref(ms)))
        return ms.value

    def records(self, run=0):
        n = int(self.problems[run].dict["scratch_size"])
        out = np.zeros(max(n, 1))
        self._check(self.lib.mbh_fetch_records(
            self.ctx, run, out.ctypes.data_as(_dp), n))
        return out[:n]

    def state(self, run=0, noalloc=False):
        """Current state of `run` (mbh_fetch_state, or with noalloc=True
        mbh_fetch_state_noalloc)."""
        d = self.problems[run].dict
        nx = int(d["num_gridpoints"])
        ds = len(d["d_init"]) // nx if nx else 0
        e, h, p = np.zeros(nx), np.zeros(nx + 1), np.zeros(nx)
        dd = np.zeros(max(nx * ds, 1))
        fn = self.lib.mbh_fetch_state_noalloc if noalloc else \
            self.lib.mbh_fetch_state
        self._check(fn(self.ctx, run, e.ctypes.data_as(_dp),
                       h.ctypes.data_as(_dp), p.ctypes.data_as(_dp),
                       dd.ctypes.data_as(_dp) if ds else None))
        return {"e": e, "h": h, "p": p, "d": dd[:nx * ds]}

    def info(self):
        i = mbh_info()
        self._check(self.lib.mbh_get_info(self.ctx, C.byref(i)))
        return {f: getattr(i, f) for f, _ in mbh_info._fields_
                if f != "reserved"}

    def close(self):
        if self.ctx:
            self.lib.mbh_destroy(self.ctx)
            self.ctx = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def split_records(problem_dict, scratch):
    """Splits a run's record scratch into {name_index: 2-D array}."""
    out = []
    for r in problem_dict["records"]:
        o, n = int(r["offset_real"]), int(r["rows"]) * int(r["cols"])
        out.append(np.asarray(scratch[o:o + n]).reshape(int(r["rows"]),
                                                        int(r["cols"])))
    return out
