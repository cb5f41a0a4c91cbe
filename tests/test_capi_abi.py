"""The C-ABI library loads and exports every symbol the header declares
(no compute calls: this runs without a GPU)."""
import ctypes as C
import os
import subprocess

import pytest

from lbfgsb_amd import _lib


def test_library_built_in_tree():
    assert os.path.exists(_lib.LIB_PATH), "run python lbfgsb-gpu_amd/build.py"
    assert "lbfgsb-gpu_amd/lib/" in _lib.LIB_PATH


def test_every_declared_symbol_is_exported():
    L = _lib.lib()
    declared = _lib.declared_symbols()
    assert len(declared) >= 20
    for name in declared:
        assert hasattr(L, name), name
    out = subprocess.run(["nm", "-D", "--defined-only", _lib.LIB_PATH],
                         capture_output=True, text=True, check=True).stdout
    exported = {ln.split()[-1] for ln in out.splitlines() if " T " in ln}
    missing = [n for n in declared if n not in exported]
    assert not missing, missing
    # every binding in _lib matches a declaration
    assert set(_lib.SIGNATURES) <= set(declared)


def test_default_options_match_reference():
    # culbfgsb/lbfgsb.cpp:44-66
    L = _lib.lib()
    o = _lib.OptionsF64()
    L.lbfgsb_default_options_f64(C.byref(o))
    assert (o.eps_f, o.eps_g, o.eps_x) == (1e-15, 1e-15, 1e-15)
    assert o.hessian_approximate_dimension == 8 and o.max_iteration == 1000
    assert o.machine_epsilon == 1e-15 and o.step_scaling == 1.0
    assert o.machine_maximum == 1.7976931348623157e308
    o32 = _lib.OptionsF32()
    L.lbfgsb_default_options_f32(C.byref(o32))
    assert abs(o32.machine_epsilon - 1e-8) < 1e-15
    assert o32.mode == 1


def test_version_string():
    assert b"gfx950" in _lib.lib().lbfgsb_amd_version()


def test_device_code_is_gfx950():
    out = subprocess.run(["readelf", "-S", _lib.LIB_PATH], capture_output=True, text=True,
                         check=True).stdout
    assert ".hip_fatbin" in out
    blob = open(_lib.LIB_PATH, "rb").read()
    assert b"gfx950" in blob


def test_struct_layouts_match_header(tmp_path):
    """The ctypes mirrors of the header's structs have the C layout (sizes
    and field offsets from the host compiler)."""
    structs = {"lbfgsb_options_f64": _lib.OptionsF64, "lbfgsb_options_f32": _lib.OptionsF32,
               "lbfgsb_summary_f64": _lib.SummaryF64, "lbfgsb_summary_f32": _lib.SummaryF32,
               "lbfgsb_callbacks_f64": _lib.CallbacksF64, "lbfgsb_request": _lib.Request,
               "lbfgsb_response": _lib.Response, "lbfgsb_serve_config": _lib.ServeConfig,
               "lbfgsb_serve_stats": _lib.ServeStats}
    lines = ['#include <stdio.h>', '#include <stddef.h>', f'#include "{_lib.HEADER}"',
             "int main(void) {"]
    for cname, py in structs.items():
        lines.append(f'  printf("{cname} size %zu\\n", sizeof({cname}));')
        for fname, _ in py._fields_:
            lines.append(f'  printf("{cname} {fname} %zu\\n", offsetof({cname}, {fname}));')
    lines += ["  return 0;", "}"]
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c99", str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout
    seen = {}
    for ln in out.splitlines():
        cname, field, val = ln.split()
        seen[(cname, field)] = int(val)
    for cname, py in structs.items():
        assert seen[(cname, "size")] == C.sizeof(py), cname
        for fname, _ in py._fields_:
            assert seen[(cname, fname)] == getattr(py, fname).offset, (cname, fname)
