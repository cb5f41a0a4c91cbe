import os
import sys

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (REPO, os.path.join(REPO, "lbfgsb-gpu_amd")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (ROCm) GPU")


def gpu_available():
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:
        return False


@pytest.fixture(scope="session")
def oracle():
    from oracle import pyoracle
    pyoracle.build()
    return pyoracle


def rel_error(f, f_ref, kind):
    """Relative objective error.  The extended Rosenbrock's minimum is
    f* = 0, so there the error is absolute (denominator 1); elsewhere it is
    |f - f_ref| / |f_ref|."""
    from lbfgsb_amd.problems import OBJ_ROSEN
    den = 1.0 if kind == OBJ_ROSEN else max(abs(f_ref), 1e-300)
    return abs(f - f_ref) / den


def solve_tolerance(eps_f, f_ref, kind, floor):
    """Bound on the relative_error of two independent solves of one problem
    that both end on the reference's decrease test: each stops once an
    iteration lowers f by at most eps_f * max(|f|, 1) (lbfgsbcpu.cpp:409-417),
    so rounding that sends them along slightly different paths can leave
    their final values a few such steps apart.  The bound is ten of them,
    or `floor`, whichever is larger."""
    from lbfgsb_amd.problems import OBJ_ROSEN
    step = 10.0 * eps_f * max(abs(f_ref), 1.0)
    den = 1.0 if kind == OBJ_ROSEN else max(abs(f_ref), 1e-300)
    return max(step / den, floor)
