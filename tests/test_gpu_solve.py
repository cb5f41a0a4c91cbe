"""Whole solves on the GPU against the CPU oracle (small sizes) -- the HIP
path through the C ABI, with the built-in objectives as funcgrad."""
import numpy as np
import pytest

from conftest import gpu_available

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not gpu_available(), reason="needs a ROCm GPU")]


def _cases():
    from lbfgsb_amd import problems as pb
    return [("grid_quadratic_64", pb.grid_quadratic(64), 1e-8, 1000),
            ("grid_cos_64", pb.grid_cos(64), 1e-7, 2000),
            ("rosenbrock_4096", pb.rosenbrock(4096), 1e-7, 2000),
            ("penta_qp_5000", pb.penta_qp(5000), 1e-6, 1000)]


@pytest.mark.parametrize("case", range(4))
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_solve_matches_oracle(case, dtype, oracle):
    from lbfgsb_amd import solver
    name, p, eps, mi = _cases()[case]
    print(f"case {name} {np.dtype(dtype).name}", flush=True)
    ref = oracle.Solver(p, dtype=np.float64, eps=eps, max_iter=mi).run()
    dp = solver.DeviceProblem.create(p, dtype)
    x, summ = solver.minimize(dp, eps=eps, max_iter=mi)
    xg = x.double().cpu().numpy()
    f_gpu, _ = __import__("lbfgsb_amd.problems", fromlist=["x"]).objective_numpy(p, xg)
    f_ref = ref.st.f
    assert summ["info"] in (1, 2, 4, 5), summ
    lo = np.where((p.nbd == 1) | (p.nbd == 2), p.l, -np.inf)
    hi = np.where((p.nbd == 2) | (p.nbd == 3), p.u, np.inf)
    slack = 1e-12 if dtype == np.float64 else 1e-6
    assert np.all(xg >= lo - slack) and np.all(xg <= hi + slack)
    rel = abs(f_gpu - f_ref) / max(abs(f_ref), 1.0)
    tol = 1e-6 if dtype == np.float64 else 1e-5
    print(name, np.dtype(dtype).name, "gpu f", f_gpu, "oracle f", f_ref, "rel", rel,
          "iters", summ["iters"], ref.st.iter, "info", summ["info"], ref.st.info)
    assert rel < tol
