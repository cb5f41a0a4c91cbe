"""ctypes binding of liblbfgsb_amd.so (the C ABI in include/lbfgsb_amd.h).

The library is loaded from the package tree (lbfgsb-gpu_amd/lib), never from
a cache or site-packages; a missing library raises, there is no fallback.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.dirname(HERE)
LIB_PATH = os.path.join(PKG, "lib", "liblbfgsb_amd.so")
HEADER = os.path.join(os.path.dirname(PKG), "include", "lbfgsb_amd.h")

_lib = None


class OptionsF64(C.Structure):
    _fields_ = [(k, C.c_double) for k in ("machine_epsilon", "machine_maximum",
                                          "step_scaling", "eps_g", "eps_f", "eps_x")] + \
               [(k, C.c_int) for k in ("hessian_approximate_dimension", "max_iteration",
                                       "mode")]


class OptionsF32(C.Structure):
    _fields_ = [(k, C.c_float) for k in ("machine_epsilon", "machine_maximum",
                                         "step_scaling", "eps_g", "eps_f", "eps_x")] + \
               [(k, C.c_int) for k in ("hessian_approximate_dimension", "max_iteration",
                                       "mode")]


class SummaryF64(C.Structure):
    _fields_ = [(k, C.c_double) for k in ("residual_g", "residual_f", "residual_x")] + \
               [("num_iteration", C.c_int), ("info", C.c_int)]


class SummaryF32(C.Structure):
    _fields_ = [(k, C.c_float) for k in ("residual_g", "residual_f", "residual_x")] + \
               [("num_iteration", C.c_int), ("info", C.c_int)]


CallbackF64 = C.CFUNCTYPE(C.c_int, C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_double),
                          C.POINTER(C.c_double), C.c_void_p, C.POINTER(SummaryF64))
CallbackF32 = C.CFUNCTYPE(C.c_int, C.c_void_p, C.POINTER(C.c_float), C.POINTER(C.c_float),
                          C.POINTER(C.c_float), C.c_void_p, C.POINTER(SummaryF32))


class CallbacksF64(C.Structure):
    _fields_ = [("funcgrad", C.c_void_p), ("funcgrad_ctx", C.c_void_p),
                ("after_iteration", C.c_void_p), ("after_iteration_ctx", C.c_void_p),
                ("customized_stopping", C.c_void_p), ("customized_stopping_ctx", C.c_void_p)]


CallbacksF32 = CallbacksF64  # same layout (function pointers and contexts)


class Request(C.Structure):
    _fields_ = [("kind", C.c_int), ("dtype", C.c_int), ("n", C.c_longlong),
                ("nx", C.c_longlong), ("ny", C.c_longlong), ("seed", C.c_ulonglong),
                ("eps", C.c_double), ("max_iteration", C.c_int), ("pad", C.c_int),
                ("arrival", C.c_double)]


class Response(C.Structure):
    _fields_ = [("f", C.c_double), ("iter", C.c_int), ("info", C.c_int), ("gang", C.c_int),
                ("dtype", C.c_int), ("t_arrival", C.c_double), ("t_admit", C.c_double),
                ("t_done", C.c_double), ("polished", C.c_int), ("iter_fp32", C.c_int)]


class ServeConfig(C.Structure):
    _fields_ = [("hbm_budget_bytes", C.c_double), ("fp32_share", C.c_double),
                ("max_slots", C.c_int), ("threads", C.c_int), ("blocks", C.c_int),
                ("segment", C.c_longlong), ("max_gang", C.c_int), ("quantum", C.c_int),
                ("use_graph", C.c_int), ("polish", C.c_int),
                ("polish_eps_scale", C.c_double)]


class ServeStats(C.Structure):
    _fields_ = [("seconds", C.c_double), ("launches", C.c_longlong),
                ("arena_bytes", C.c_double), ("max_waiting", C.c_int),
                ("max_resident", C.c_int)]

# Every exported symbol with its signature (restype, argtypes).
SIGNATURES = {
    "lbfgsb_default_options_f64": (None, [C.POINTER(OptionsF64)]),
    "lbfgsb_default_options_f32": (None, [C.POINTER(OptionsF32)]),
    "lbfgsb_minimize_f64": (C.c_int, [C.c_int, C.POINTER(CallbacksF64), C.POINTER(OptionsF64),
                                      C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                      C.POINTER(SummaryF64), C.c_void_p]),
    "lbfgsb_minimize_f32": (C.c_int, [C.c_int, C.POINTER(CallbacksF32), C.POINTER(OptionsF32),
                                      C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                      C.POINTER(SummaryF32), C.c_void_p]),
    "lbfgsb_minimize_noalloc_f64": (C.c_int, [C.c_int, C.POINTER(CallbacksF64),
                                              C.POINTER(OptionsF64), C.c_void_p, C.c_void_p,
                                              C.c_void_p, C.c_void_p, C.POINTER(SummaryF64),
                                              C.c_void_p]),
    "lbfgsb_minimize_noalloc_f32": (C.c_int, [C.c_int, C.POINTER(CallbacksF32),
                                              C.POINTER(OptionsF32), C.c_void_p, C.c_void_p,
                                              C.c_void_p, C.c_void_p, C.POINTER(SummaryF32),
                                              C.c_void_p]),
    "lbfgsb_engine_create": (C.c_void_p, [C.c_int, C.c_longlong, C.c_int, C.c_void_p,
                                          C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    "lbfgsb_engine_destroy": (None, [C.c_void_p]),
    "lbfgsb_engine_setup_begin": (C.c_int, [C.c_void_p, C.c_double, C.c_double, C.c_double,
                                            C.c_int]),
    "lbfgsb_engine_setup_finish": (C.c_int, [C.c_void_p, C.c_double]),
    "lbfgsb_engine_run": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_double]),
    "lbfgsb_engine_x": (C.c_void_p, [C.c_void_p, C.c_int]),
    "lbfgsb_engine_g": (C.c_void_p, [C.c_void_p, C.c_int]),
    "lbfgsb_engine_get_scalar": (C.c_int, [C.c_void_p, C.c_char_p, C.POINTER(C.c_double)]),
    "lbfgsb_engine_set_scalar": (C.c_int, [C.c_void_p, C.c_char_p, C.c_double]),
    "lbfgsb_engine_get_array": (C.c_int, [C.c_void_p, C.c_char_p, C.c_void_p, C.c_longlong]),
    "lbfgsb_engine_set_array": (C.c_int, [C.c_void_p, C.c_char_p, C.c_void_p, C.c_longlong]),
    "lbfgsb_engine_finalize": (C.c_int, [C.c_void_p]),
    "lbfgsb_engine_solve_f64": (C.c_int, [C.c_void_p, C.POINTER(CallbacksF64),
                                          C.POINTER(OptionsF64), C.POINTER(SummaryF64),
                                          C.c_void_p, C.c_int]),
    "lbfgsb_engine_solve_f32": (C.c_int, [C.c_void_p, C.POINTER(CallbacksF32),
                                          C.POINTER(OptionsF32), C.POINTER(SummaryF32),
                                          C.c_void_p, C.c_int]),
    "lbfgsb_engine_set_profiling": (None, [C.c_void_p, C.c_int]),
    "lbfgsb_engine_set_profiling_steps": (None, [C.c_void_p, C.c_int, C.c_uint]),
    "lbfgsb_engine_step_time": (C.c_int, [C.c_void_p, C.c_int, C.POINTER(C.c_double),
                                          C.POINTER(C.c_longlong)]),
    "lbfgsb_problem_create": (C.c_void_p, [C.c_int, C.c_int, C.c_longlong, C.c_longlong,
                                           C.c_longlong, C.c_double, C.c_void_p, C.c_void_p,
                                           C.c_void_p, C.c_void_p, C.c_void_p]),
    "lbfgsb_problem_destroy": (None, [C.c_void_p]),
    "lbfgsb_problem_funcgrad_f64": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                              C.c_void_p, C.c_void_p]),
    "lbfgsb_problem_funcgrad_f32": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                              C.c_void_p, C.c_void_p]),
    "lbfgsb_problem_eval_device": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                             C.c_void_p]),
    "lbfgsb_batch_create": (C.c_void_p, [C.c_int, C.c_longlong, C.c_int, C.c_int, C.c_void_p]),
    "lbfgsb_batch_destroy": (None, [C.c_void_p]),
    "lbfgsb_batch_add": (C.c_int, [C.c_void_p, C.c_longlong, C.c_void_p, C.c_void_p, C.c_void_p,
                                   C.c_void_p, C.c_int, C.c_longlong, C.c_longlong, C.c_double,
                                   C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                   C.c_double, C.c_double, C.c_double, C.c_int]),
    "lbfgsb_batch_run": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_int]),
    "lbfgsb_batch_result": (C.c_int, [C.c_void_p, C.c_int, C.POINTER(C.c_double),
                                      C.POINTER(C.c_int), C.POINTER(C.c_int),
                                      C.POINTER(C.c_double), C.POINTER(C.c_int)]),
    "lbfgsb_batch_results": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                       C.c_void_p]),
    "lbfgsb_batch_clocks": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p,
                                      C.POINTER(C.c_double)]),
    "lbfgsb_batch_get_x": (C.c_int, [C.c_void_p, C.c_int, C.c_void_p]),
    "lbfgsb_batch_warm_start": (C.c_int, [C.c_void_p, C.c_int, C.c_void_p, C.c_int]),
    "lbfgsb_batch_remove": (C.c_int, [C.c_void_p, C.c_int]),
    "lbfgsb_batch_free_elems": (C.c_longlong, [C.c_void_p]),
    "lbfgsb_batch_configure": (C.c_int, [C.c_void_p, C.c_longlong, C.c_int]),
    "lbfgsb_batch_gang": (C.c_int, [C.c_void_p, C.c_int]),
    "lbfgsb_batch_use_graph": (C.c_int, [C.c_void_p, C.c_int]),
    "lbfgsb_batch_bytes": (C.c_longlong, [C.c_int, C.c_longlong, C.c_int]),
    "lbfgsb_batch_add_generated": (C.c_int, [C.c_void_p, C.c_int, C.c_longlong, C.c_longlong,
                                             C.c_longlong, C.c_ulonglong, C.c_double,
                                             C.c_double, C.c_double, C.c_int]),
    "lbfgsb_generate": (C.c_int, [C.c_int, C.c_int, C.c_longlong, C.c_longlong, C.c_longlong,
                                  C.c_ulonglong] + [C.c_void_p] * 10),
    "lbfgsb_serve_stream": (C.c_int, [C.POINTER(ServeConfig), C.POINTER(Request), C.c_int,
                                      C.POINTER(Response), C.POINTER(ServeStats), C.c_void_p]),
    "lbfgsb_amd_version": (C.c_char_p, []),
}


def lib():
    """Load the in-tree library (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"{LIB_PATH} is missing: build it with `python lbfgsb-gpu_amd/build.py`")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def declared_symbols(header: str = HEADER):
    """Function names declared in include/lbfgsb_amd.h."""
    import re
    text = open(header).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"\b(lbfgsb_\w+)\s*\(", text)
    return sorted(set(n for n in names if not n.startswith("lbfgsb_callback")))
