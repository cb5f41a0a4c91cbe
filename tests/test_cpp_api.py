"""The C++ drop-in API (include/culbfgsb.h) and its example program.

The header restates the reference's culbfgsb/culbfgsb.h types and the
lbfgsbcuda::lbfgsbminimize / lbfgsbdefaultoption templates over the C ABI;
examples/dsscfg.hip is the reference's example on the reference's problem
(MINPACK-2 dsscfg, examples/dsscfg/dsscfg_main.cpp): a user callback with its
own HIP kernels on the solve's stream, float and double solves.
"""
import os
import subprocess

import pytest

from conftest import REPO, gpu_available

PKG = os.path.join(REPO, "lbfgsb-gpu_amd")
EXAMPLE = os.path.join(PKG, "lib", "dsscfg_example")

_TU = r"""
#include <cstring>
#include "culbfgsb.h"
int use_api(int n, float* x, const int* nbd, const float* l, const float* u) {
  LBFGSB_CUDA_OPTION<float> opt;
  lbfgsbcuda::lbfgsbdefaultoption<float>(opt);
  LBFGSB_CUDA_STATE<float> state;
  std::memset(&state, 0, sizeof(state));
  state.m_funcgrad_callback = [](float* x, float& f, float* g, const hipStream_t& s,
                                 const LBFGSB_CUDA_SUMMARY<float>& sm) { f = 0; return 0; };
  LBFGSB_CUDA_SUMMARY<float> summary;
  lbfgsbcuda::lbfgsbminimize<float>(n, state, opt, x, nbd, l, u, summary);
  LBFGSB_CUDA_OPTION<double> o2;
  lbfgsbcuda::lbfgsbdefaultoption<double>(o2);
  static_assert(sizeof(LBFGSB_CUDA_SUMMARY<double>) == sizeof(lbfgsb_summary_f64), "layout");
  return summary.info + (o2.mode == LCM_CUDA ? 0 : 1);
}
"""


def test_header_compiles_with_host_compiler(tmp_path):
    """A host-only translation unit (g++, no HIP compiler) that uses every
    name of the reference API compiles against the header."""
    src = tmp_path / "use_api.cpp"
    src.write_text(_TU)
    cmd = ["g++", "-std=c++17", "-fsyntax-only", "-D__HIP_PLATFORM_AMD__",
           "-I" + os.path.join(REPO, "include"), "-I/opt/rocm/include", str(src)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


def test_example_is_built():
    assert os.path.exists(EXAMPLE), "run __graft_entry__.build() first"


@pytest.mark.gpu
@pytest.mark.skipif(not gpu_available(), reason="needs a ROCm GPU")
def test_example_runs(oracle):
    """The reference's only test, on the reference's problem: MINPACK-2
    dsscfg, 256 x 256, lambda = 1, unconstrained, eps 1e-8, 1000 iterations
    (examples/dsscfg/dsscfg_main.cpp:21-24, 35-40, 69-80).  As there, the
    least f the callback sees must agree with the CPU path's within 1e-4 in
    double and 1e-3 in float (:208-231); the CPU path is the oracle.  The
    example itself also checks both precisions converge (|g|_inf, info,
    one after-iteration call per iteration) and agree with each other."""
    import re

    import numpy as np
    from lbfgsb_amd import problems as pb
    print("case dsscfg_example 256 x 256", flush=True)
    r = subprocess.run([EXAMPLE], capture_output=True, text=True, timeout=300)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.strip().endswith("OK")
    m = re.search(r"min f: double (\S+), float (\S+),", r.stdout)
    f64, f32 = float(m.group(1)), float(m.group(2))
    p = pb.dsscfg(256)
    for dt, f, tol in ((np.float64, f64, 1e-4), (np.float32, f32, 1e-3)):
        ref = oracle.Solver(p, dtype=dt, eps=1e-8, max_iter=1000).run()
        f_cpu = min(rec.f for rec in ref.records)
        print(np.dtype(dt).name, "example", f, "oracle", f_cpu, "iterations", ref.st.iter)
        assert abs(f - f_cpu) < tol
