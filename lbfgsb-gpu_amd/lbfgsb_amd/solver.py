"""Python host layer over the C ABI: device problems, the plain solve and the
stepping engine.  PyTorch only provides device memory and the stream."""
from __future__ import annotations

import ctypes as C
import dataclasses

import numpy as np

from . import _lib
from .problems import Problem

# engine status codes (DevScalars::status)
(ST_RUNNING, ST_NEED_EVAL, ST_ACCEPTED, ST_CHECKED, ST_GCP_ROUNDS, ST_RESTART, ST_DONE,
 ST_REDO) = range(8)
# programs and steps of lbfgsb_engine_run
PROG_ITER_START, PROG_AFTER_EVAL, PROG_AFTER_ACCEPT, PROG_AFTER_CHECK, PROG_TAIL = range(5)
(K_RESUME_RESET, K_RESUME, K_PASSA_SA, K_UPD_REDUCE, K_PASSA_FUSED, K_STOP, K_UPD, K_P4,
 K_FORMK, K_P5A, K_ALPHA, K_P5B, K_LSINIT, K_TRIAL, K_GD, K_LSSTEP, K_END) = range(17)


def _torch():
    import torch
    if not torch.cuda.is_available():
        raise RuntimeError("lbfgsb_amd needs a ROCm GPU (torch.cuda.is_available() is False)")
    return torch


def _stream_ptr(torch):
    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


@dataclasses.dataclass
class DeviceProblem:
    """A Problem resident in HBM plus its built-in objective handle."""
    problem: Problem
    dtype: type
    tensors: dict
    handle: int

    @classmethod
    def create(cls, p: Problem, dtype=np.float64):
        torch = _torch()
        tdt = torch.float64 if np.dtype(dtype) == np.float64 else torch.float32
        q = p.cast(dtype)
        t = {"x0": torch.from_numpy(q.x0).to("cuda", tdt),
             "l": torch.from_numpy(q.l).to("cuda", tdt),
             "u": torch.from_numpy(q.u).to("cuda", tdt),
             "nbd": torch.from_numpy(q.nbd).to("cuda", torch.int32)}
        for k in ("b", "c", "a0", "a1", "a2"):
            a = getattr(q, k)
            if a is not None:
                t[k] = torch.from_numpy(a).to("cuda", tdt)
        ptr = lambda k: C.c_void_p(t[k].data_ptr()) if k in t else None
        h = _lib.lib().lbfgsb_problem_create(
            1 if tdt == torch.float64 else 0, p.kind, p.n, p.nx, p.ny, p.h,
            ptr("b"), ptr("c"), ptr("a0"), ptr("a1"), ptr("a2"))
        if not h:
            raise RuntimeError("lbfgsb_problem_create failed")
        return cls(p, dtype, t, h)

    def close(self):
        if self.handle:
            _lib.lib().lbfgsb_problem_destroy(self.handle)
            self.handle = 0

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def f64(self):
        return np.dtype(self.dtype) == np.float64


def minimize(dp: DeviceProblem, eps=1e-8, max_iter=1000, m=8, x0=None, eps_g=None,
             eps_f=None, eps_x=None, after_iteration=None, stop=None, noalloc=False,
             foreign_funcgrad=False):
    """Plain solve through lbfgsb_minimize_{f64,f32} (noalloc: through
    lbfgsb_minimize_noalloc_{f64,f32}) with the built-in objective as
    funcgrad.  Returns (x tensor, summary dict with the return code).

    The library recognises its own objective and takes each trial's value
    from device memory; with foreign_funcgrad the objective is passed inside
    a ctypes callback instead, which the library treats as a caller's
    funcgrad (the value returned on the host after each evaluation)."""
    torch = _torch()
    L = _lib.lib()
    x = (dp.tensors["x0"] if x0 is None else x0).clone()
    f64 = dp.f64
    opt = _lib.OptionsF64() if f64 else _lib.OptionsF32()
    (L.lbfgsb_default_options_f64 if f64 else L.lbfgsb_default_options_f32)(C.byref(opt))
    opt.eps_g = eps if eps_g is None else eps_g
    opt.eps_f = eps if eps_f is None else eps_f
    opt.eps_x = eps if eps_x is None else eps_x
    opt.max_iteration = max_iter
    opt.hessian_approximate_dimension = m
    cbs = _lib.CallbacksF64()
    fg = L.lbfgsb_problem_funcgrad_f64 if f64 else L.lbfgsb_problem_funcgrad_f32
    cbs.funcgrad = C.cast(fg, C.c_void_p)
    cbs.funcgrad_ctx = C.c_void_p(dp.handle)
    keep = []
    CB = _lib.CallbackF64 if f64 else _lib.CallbackF32
    if foreign_funcgrad:
        wrap = CB(lambda ctx, xp, fp, gp, s, summ: fg(ctx, C.cast(xp, C.c_void_p),
                                                      C.cast(fp, C.c_void_p),
                                                      C.cast(gp, C.c_void_p), s,
                                                      C.cast(summ, C.c_void_p)))
        keep.append(wrap)
        cbs.funcgrad = C.cast(wrap, C.c_void_p)
    if after_iteration is not None:
        cb = CB(lambda ctx, xp, fp, gp, s, summ: int(after_iteration(summ.contents) or 0))
        keep.append(cb)
        cbs.after_iteration = C.cast(cb, C.c_void_p)
    if stop is not None:
        cb2 = CB(lambda ctx, xp, fp, gp, s, summ: int(bool(stop(summ.contents))))
        keep.append(cb2)
        cbs.customized_stopping = C.cast(cb2, C.c_void_p)
    summ = _lib.SummaryF64() if f64 else _lib.SummaryF32()
    if noalloc:
        fn = L.lbfgsb_minimize_noalloc_f64 if f64 else L.lbfgsb_minimize_noalloc_f32
    else:
        fn = L.lbfgsb_minimize_f64 if f64 else L.lbfgsb_minimize_f32
    t = dp.tensors
    rc = fn(dp.problem.n, C.byref(cbs), C.byref(opt), C.c_void_p(x.data_ptr()),
            C.c_void_p(t["nbd"].data_ptr()), C.c_void_p(t["l"].data_ptr()),
            C.c_void_p(t["u"].data_ptr()), C.byref(summ), _stream_ptr(torch))
    torch.cuda.synchronize()
    return x, dict(info=summ.info, iters=summ.num_iteration, residual_g=summ.residual_g,
                   residual_f=summ.residual_f, residual_x=summ.residual_x, rc=rc)


class Engine:
    """Stepping view of the device engine (lbfgsb_engine_*)."""

    def __init__(self, dp: DeviceProblem, m=8, x0=None):
        torch = _torch()
        self.torch = torch
        self.dp = dp
        self.L = _lib.lib()
        self.n = dp.problem.n
        self.m = m
        self.np_dtype = np.float64 if dp.f64 else np.float32
        self.x = (dp.tensors["x0"] if x0 is None else x0).clone()
        t = dp.tensors
        self.stream = _stream_ptr(torch)
        self.h = self.L.lbfgsb_engine_create(
            1 if dp.f64 else 0, self.n, m, C.c_void_p(self.x.data_ptr()),
            C.c_void_p(t["nbd"].data_ptr()), C.c_void_p(t["l"].data_ptr()),
            C.c_void_p(t["u"].data_ptr()), self.stream)
        if not self.h:
            raise RuntimeError("lbfgsb_engine_create failed")
        self.f = 0.0
        self.evals = 0

    def close(self):
        if self.h:
            self.L.lbfgsb_engine_destroy(self.h)
            self.h = 0

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def evaluate(self, which):
        """Built-in objective at x_which into g_which; returns f."""
        xp = self.L.lbfgsb_engine_x(self.h, which)
        gp = self.L.lbfgsb_engine_g(self.h, which)
        f = C.c_double(0) if self.dp.f64 else C.c_float(0)
        fg = (self.L.lbfgsb_problem_funcgrad_f64 if self.dp.f64
              else self.L.lbfgsb_problem_funcgrad_f32)
        fg(C.c_void_p(self.dp.handle), C.c_void_p(xp), C.byref(f), C.c_void_p(gp),
           self.stream, None)
        self.evals += 1
        return float(f.value)

    def setup(self, eps=1e-8, max_iter=1000, eps_g=None, eps_f=None, eps_x=None):
        st = self.L.lbfgsb_engine_setup_begin(
            self.h, eps if eps_g is None else eps_g, eps if eps_f is None else eps_f,
            eps if eps_x is None else eps_x, max_iter)
        if st == ST_DONE:
            return st
        self.f = self.evaluate(0)
        return self.L.lbfgsb_engine_setup_finish(self.h, self.f)

    def run(self, program, start=-1, stop=K_END, f=0.0):
        return self.L.lbfgsb_engine_run(self.h, program, start, stop, f)

    def scalar(self, name):
        v = C.c_double(0)
        if self.L.lbfgsb_engine_get_scalar(self.h, name.encode(), C.byref(v)) != 0:
            raise KeyError(name)
        return v.value

    def set_scalar(self, name, value):
        if self.L.lbfgsb_engine_set_scalar(self.h, name.encode(), float(value)) != 0:
            raise KeyError(name)

    def array(self, name, count=None):
        n, m = self.n, self.m
        if name in ("iwhere", "index", "indx2"):
            out = np.zeros(n, dtype=np.int32)
        elif name == "iwbits":
            out = np.zeros(n, dtype=np.uint8)
        elif name in ("ws", "wy"):
            out = np.zeros(n * m, dtype=self.np_dtype)
        elif name in ("sy", "ss", "wt"):
            out = np.zeros(m * m)
        elif name in ("wn", "snd"):
            out = np.zeros(4 * m * m)
        elif name in ("p", "c", "mc", "wv"):
            out = np.zeros(16 if count is None else count)
        elif name == "pa":
            out = np.zeros(39)
        elif name == "dsave":
            out = np.zeros(13)
        elif name == "isave":
            out = np.zeros(2)
        else:
            out = np.zeros(n, dtype=self.np_dtype)
        rc = self.L.lbfgsb_engine_get_array(self.h, name.encode(),
                                            out.ctypes.data_as(C.c_void_p), out.size)
        if rc != 0:
            raise RuntimeError(f"get_array({name}) -> {rc}")
        return out

    def set_array(self, name, a):
        a = np.ascontiguousarray(a)
        rc = self.L.lbfgsb_engine_set_array(self.h, name.encode(),
                                            a.ctypes.data_as(C.c_void_p), a.size)
        if rc != 0:
            raise RuntimeError(f"set_array({name}) -> {rc}")

    def solve(self, eps=1e-8, max_iter=1000, max_evals=100000):
        """Drive a whole solve with the built-in objective (host loop of
        lbfgsb_minimize, exposed step by step)."""
        st = self.setup(eps, max_iter)
        prog, f = PROG_ITER_START, 0.0
        while st not in (ST_DONE, -1) and self.evals < max_evals:
            st = self.run(prog, f=f)
            if st == ST_NEED_EVAL:
                f = self.evaluate(1)
                prog = PROG_AFTER_EVAL
            elif st == ST_RESTART:
                prog = PROG_ITER_START
            elif st in (ST_ACCEPTED, ST_REDO):
                prog = PROG_AFTER_ACCEPT
            elif st == ST_CHECKED:
                prog = PROG_AFTER_CHECK
        self.L.lbfgsb_engine_finalize(self.h)
        return st
