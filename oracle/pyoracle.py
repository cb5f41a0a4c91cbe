"""ctypes view of the CPU oracle (liblbfgsb_oracle.so).  TEST INFRASTRUCTURE.

Only tests, the smoke check and the benchmark's CPU-baseline leg use this
module.  It exposes the oracle's whole-solve entry, its per-iteration
stepping and every component function on an explicit state, with numpy
views over the state arrays.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liblbfgsb_oracle.so")

OBJ_NONE, OBJ_GRID_QUAD, OBJ_GRID_COS, OBJ_ROSEN, OBJ_PENTA, OBJ_USER, OBJ_DSSCFG = range(7)
TAIL_NEXT, TAIL_RESTART, TAIL_END = range(3)


def build(force: bool = False) -> str:
    if force or not os.path.exists(LIB_PATH):
        subprocess.run(["make", "-s", "-C", HERE], check=True)
    return LIB_PATH


def _types(real):
    P = C.POINTER(real)
    PI = C.POINTER(C.c_int)
    FG = C.CFUNCTYPE(C.c_int, C.c_void_p, P, P, P)

    class Objective(C.Structure):
        _fields_ = [("kind", C.c_int), ("nx", C.c_int), ("ny", C.c_int),
                    ("h", C.c_double), ("b", P), ("c", P), ("a0", P),
                    ("a1", P), ("a2", P), ("user", FG), ("user_ctx", C.c_void_p)]

    class Options(C.Structure):
        _fields_ = [("eps_g", C.c_double), ("eps_f", C.c_double),
                    ("eps_x", C.c_double), ("m", C.c_int),
                    ("max_iteration", C.c_int), ("stop_at_iter", C.c_int),
                    ("pad_", C.c_int)]

    class Record(C.Structure):
        _fields_ = [(k, C.c_double) for k in
                    ("f", "sbgnrm", "stp", "theta", "dtd", "gd", "residual_f",
                     "residual_x")] + \
                   [(k, C.c_int) for k in
                    ("iter", "nint", "nfree", "nenter", "ileave", "wrk",
                     "iword", "iback", "nfgv", "col", "updatd", "restarts",
                     "info", "pad")]

    int_scalars = ["n", "m", "col", "head", "itail", "iupdat", "iter", "nfgv",
                   "nint", "nintol", "nskip", "nfree", "nact", "ileave",
                   "nenter", "iword", "ifun", "iback", "info", "task", "csave",
                   "internalinfo", "prjctd", "cnstnd", "boxed", "updatd", "wrk",
                   "k_err", "restarts"]
    real_scalars = ["f", "fold", "theta", "dr", "rr", "dnrm", "xstep", "sbgnrm",
                    "ddum", "dtd", "gd", "gdold", "stp", "stpmx", "tf",
                    "residual_g", "residual_f", "residual_x"]

    class State(C.Structure):
        _fields_ = ([(k, C.c_int) for k in int_scalars] +
                    [("isave", C.c_int * 2)] +
                    [(k, real) for k in real_scalars] +
                    [("dsave", real * 13), ("wa", real * 64)] +
                    [(k, P) for k in ("x", "g", "l", "u", "xold", "xdiff", "z",
                                      "r", "d", "t", "ws", "wy", "sy", "ss",
                                      "wt", "wn", "snd")] +
                    [(k, PI) for k in ("nbd", "index", "iwhere", "indx2")] +
                    [("obj", Objective), ("opt", Options)])

    return Objective, Options, Record, State, FG


class _Flavor:
    def __init__(self, lib, real, sfx):
        self.lib, self.real, self.sfx = lib, real, sfx
        self.np = np.float64 if real is C.c_double else np.float32
        (self.Objective, self.Options, self.Record, self.State,
         self.FG) = _types(real)
        P = C.POINTER(real)
        PI = C.POINTER(C.c_int)
        PS = C.POINTER(self.State)

        def fn(name, res, *args):
            f = getattr(lib, f"{name}_{sfx}")
            f.restype = res
            f.argtypes = list(args)
            return f

        self.create = fn("orc_create", PS, C.c_int, C.c_int)
        self.destroy = fn("orc_destroy", None, PS)
        self.copy = fn("orc_copy", None, PS, PS)
        self.evaluate = fn("orc_evaluate", C.c_int, C.POINTER(self.Objective),
                           C.c_int, P, P, P)
        self.setup = fn("orc_setup", C.c_int, PS, P, PI, P, P,
                        C.POINTER(self.Options), C.POINTER(self.Objective))
        self.iterate = fn("orc_iterate", C.c_int, PS, C.POINTER(self.Record))
        self.iter_head = fn("orc_iter_head", None, PS)
        self.iter_search = fn("orc_iter_search", None, PS)
        self.iter_tail = fn("orc_iter_tail", C.c_int, PS)
        self.lnsrlb_step = fn("orc_lnsrlb_step", C.c_int, PS)
        self.minimize = fn("orc_minimize", C.c_int, C.c_int,
                           C.POINTER(self.Options), C.POINTER(self.Objective),
                           P, PI, P, P, C.POINTER(self.Record), C.c_int,
                           PI, P)
        for comp in ("cauchy", "freev", "formk", "cmprlb", "subsm", "matupd",
                     "formt", "projgr", "active"):
            setattr(self, comp, fn(f"orc_{comp}", None, PS))
        self.dpofa = fn("orc_dpofa", C.c_int, P, C.c_int, C.c_int)
        self.dtrsl = fn("orc_dtrsl", C.c_int, P, C.c_int, C.c_int, P, C.c_int)
        self.bmv = fn("orc_bmv", C.c_int, C.c_int, P, P, C.c_int, P, P)
        self.hpsolb = fn("orc_hpsolb", None, C.c_int, P, PI, C.c_int)
        self.dcsrch = fn("orc_dcsrch", None, real, real, P, real, real, real,
                         real, real, PI, PI, P)
        self.dcstep = fn("orc_dcstep", None, P, P, P, P, P, P, P, real, real,
                         PI, real, real)
        self.dot = fn("orc_dot", real, P, C.c_int, P, C.c_int, C.c_int)

    def ptr(self, a):
        return a.ctypes.data_as(C.POINTER(self.real))


_LIB = None
F64 = F32 = None


def load():
    global _LIB, F64, F32
    if _LIB is None:
        build()
        _LIB = C.CDLL(LIB_PATH)
        F64 = _Flavor(_LIB, C.c_double, "f64")
        F32 = _Flavor(_LIB, C.c_float, "f32")
    return F64, F32


def flavor(dtype):
    f64, f32 = load()
    return f64 if np.dtype(dtype) == np.float64 else f32


def _iptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_int))


class Objective:
    """Keeps numpy arrays alive for the C objective struct."""

    def __init__(self, fl, problem):
        self.fl = fl
        dt = fl.np
        self.keep = []

        def arr(a):
            if a is None:
                return None
            a = np.ascontiguousarray(a, dtype=dt)
            self.keep.append(a)
            return fl.ptr(a)

        self.c = fl.Objective()
        self.c.kind = problem.kind
        self.c.nx, self.c.ny, self.c.h = problem.nx, problem.ny, problem.h
        for k in ("b", "c", "a0", "a1", "a2"):
            p = arr(getattr(problem, k))
            if p is not None:
                setattr(self.c, k, p)


class Solver:
    """Stepping view of one oracle solve."""

    def __init__(self, problem, dtype=np.float64, m=8, eps=1e-8, max_iter=1000,
                 eps_g=None, eps_f=None, eps_x=None, stop_at_iter=0):
        self.fl = fl = flavor(dtype)
        self.p = problem.cast(fl.np)
        self.n = n = problem.n
        self.obj = Objective(fl, self.p)
        self.opt = fl.Options(eps if eps_g is None else eps_g,
                              eps if eps_f is None else eps_f,
                              eps if eps_x is None else eps_x, m, max_iter,
                              stop_at_iter, 0)
        self.s = fl.create(n, m)
        self.more = bool(fl.setup(self.s, fl.ptr(self.p.x0), _iptr(self.p.nbd),
                                  fl.ptr(self.p.l), fl.ptr(self.p.u),
                                  C.byref(self.opt), C.byref(self.obj.c)))
        self.records = []

    def __del__(self):
        try:
            self.fl.destroy(self.s)
        except Exception:
            pass

    @property
    def st(self):
        return self.s.contents

    def arr(self, name, count=None):
        st = self.st
        n, m = st.n, st.m
        sizes = {"ws": n * m, "wy": n * m, "sy": m * m, "ss": m * m,
                 "wt": m * m, "wn": 4 * m * m, "snd": 4 * m * m}
        cnt = count if count is not None else sizes.get(name, n)
        ptr = getattr(st, name)
        return np.ctypeslib.as_array(ptr, shape=(cnt,))

    def step(self):
        rec = self.fl.Record()
        self.more = bool(self.fl.iterate(self.s, C.byref(rec)))
        self.records.append(rec)
        return rec

    def run(self, max_steps=10**9):
        while self.more and len(self.records) < max_steps:
            self.step()
        return self

    def clone(self):
        other = object.__new__(Solver)
        other.__dict__.update(self.__dict__)
        other.s = self.fl.create(self.n, self.st.m)
        self.fl.copy(other.s, self.s)
        other.records = list(self.records)
        return other


def minimize(problem, dtype=np.float64, m=8, eps=1e-8, max_iter=1000,
             max_records=100000, eps_g=None, eps_f=None, eps_x=None, stop_at_iter=0):
    fl = flavor(dtype)
    p = problem.cast(fl.np)
    obj = Objective(fl, p)
    opt = fl.Options(eps if eps_g is None else eps_g, eps if eps_f is None else eps_f,
                     eps if eps_x is None else eps_x, m, max_iter, stop_at_iter, 0)
    x = p.x0.copy()
    recs = (fl.Record * max_records)()
    iters = C.c_int(0)
    fval = fl.real(0)
    info = fl.minimize(problem.n, C.byref(opt), C.byref(obj.c), fl.ptr(x),
                       _iptr(p.nbd), fl.ptr(p.l), fl.ptr(p.u), recs,
                       max_records, C.byref(iters), C.byref(fval))
    nrec = min(iters.value + 1, max_records)
    return dict(info=info, x=x, f=float(fval.value), iters=iters.value,
                records=[recs[i] for i in range(nrec)])
