"""The CMake package (lbfgsb-gpu_amd/CMakeLists.txt): configure, build and
install the library with its headers, then build a consumer that finds it
with find_package(lbfgsb_amd) and calls the reference's C++ API (no GPU is
needed: the consumer is compiled and linked, not run)."""
import os
import shutil
import subprocess

import pytest

from conftest import REPO

ROCM = "/opt/rocm"
CLANG = os.path.join(ROCM, "lib", "llvm", "bin", "clang++")

pytestmark = pytest.mark.skipif(not (shutil.which("cmake") and os.path.exists(CLANG)),
                                reason="needs cmake and the ROCm clang")

_CONSUMER = r"""
#include <cstring>
#include "culbfgsb.h"
int main() {
  LBFGSB_CUDA_OPTION<double> opt;
  lbfgsbcuda::lbfgsbdefaultoption<double>(opt);
  LBFGSB_CUDA_STATE<double> state;
  std::memset(&state, 0, sizeof(state));
  return opt.max_iteration == 1000 && opt.mode == LCM_CUDA ? 0 : 1;
}
"""

_CONSUMER_CMAKE = """
cmake_minimum_required(VERSION 3.21)
project(consumer LANGUAGES CXX)
set(CMAKE_CXX_STANDARD 17)
find_package(lbfgsb_amd REQUIRED)
add_executable(consumer main.cpp)
target_link_libraries(consumer PRIVATE lbfgsb_amd::lbfgsb_amd)
"""


def _run(cmd, **kw):
    env = dict(os.environ, PATH=os.path.join(ROCM, "bin") + ":" + os.environ.get("PATH", ""))
    r = subprocess.run(cmd, capture_output=True, text=True, env=env, **kw)
    assert r.returncode == 0, (cmd, r.stdout[-2000:], r.stderr[-2000:])
    return r


def test_cmake_build_install_and_consume(tmp_path):
    build, inst = tmp_path / "build", tmp_path / "inst"
    common = [f"-DCMAKE_HIP_COMPILER={CLANG}", f"-DCMAKE_CXX_COMPILER={CLANG}",
              f"-DCMAKE_PREFIX_PATH={ROCM};{inst}"]
    _run(["cmake", "-S", os.path.join(REPO, "lbfgsb-gpu_amd"), "-B", str(build),
          f"-DCMAKE_INSTALL_PREFIX={inst}"] + common)
    _run(["cmake", "--build", str(build), "-j", str(min(8, os.cpu_count() or 1))], timeout=1200)
    _run(["cmake", "--install", str(build)])
    for f in ("include/culbfgsb.h", "include/lbfgsb_amd.h", "lib/liblbfgsb_amd.so",
              "lib/cmake/lbfgsb_amd/lbfgsb_amdConfig.cmake"):
        assert (inst / f).exists(), f
    syms = _run(["nm", "-D", "--defined-only", str(inst / "lib" / "liblbfgsb_amd.so")]).stdout
    for name in ("lbfgsb_minimize_f64", "lbfgsb_minimize_f32", "lbfgsb_default_options_f64",
                 "lbfgsb_minimize_noalloc_f64"):
        assert name in syms, name
    # built for gfx950 when the configure names no architecture (the device
    # code objects carry their target name)
    blob = (inst / "lib" / "liblbfgsb_amd.so").read_bytes()
    assert b"gfx950" in blob and b"gfx906" not in blob
    src = tmp_path / "consumer"
    src.mkdir()
    (src / "main.cpp").write_text(_CONSUMER)
    (src / "CMakeLists.txt").write_text(_CONSUMER_CMAKE)
    cb = tmp_path / "cbuild"
    _run(["cmake", "-S", str(src), "-B", str(cb)] + common)
    _run(["cmake", "--build", str(cb)])
    assert (cb / "consumer").exists()
