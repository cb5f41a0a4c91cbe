"""Whole solves on the GPU against the CPU oracle (small sizes) -- the HIP
path through the C ABI, with the built-in objectives as funcgrad.

fp64: the same info as the oracle.  The two evaluate f and g with
different summation orders (and the device contracts products into FMAs),
so trajectories agree to rounding only as far as the problem keeps rounding
from growing: on the convex quadratics and the cos-energy grid the
iteration count is equal and the final f agrees to 1e-12 relative; on the
extended Rosenbrock and the dsscfg energy the count may differ by 2 and the
final f by 10 eps_f (the resolution of the decrease test that ends both
solves).  Decisions on identical inputs are compared exactly by the
lockstep tests.  fp32 (FP32 vectors, FP64
scalar logic) against the FP64 oracle: the final f within the relative
bound stated per case (the float iterate's own resolution), info a
convergence code.  Relative error is |f - f_ref| / |f_ref|, absolute for the
extended Rosenbrock (f* = 0)."""
import numpy as np
import pytest

from conftest import gpu_available, rel_error, solve_tolerance

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not gpu_available(), reason="needs a ROCm GPU")]


def _cases():
    from lbfgsb_amd import problems as pb
    # name, problem, eps, max_iteration, fp32 bound on the relative f error,
    # fp64 trajectory equal to the oracle's
    return [("grid_quadratic_64", pb.grid_quadratic(64), 1e-8, 1000, 2e-5, True),
            ("grid_cos_64", pb.grid_cos(64), 1e-7, 2000, 1e-5, True),
            ("rosenbrock_4096", pb.rosenbrock(4096), 1e-7, 2000, 1e-5, False),
            ("penta_qp_5000", pb.penta_qp(5000), 1e-6, 1000, 1e-6, True),
            ("dsscfg_64", pb.dsscfg(64), 1e-8, 1000, 1e-6, False)]


@pytest.mark.parametrize("case", range(5))
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_solve_matches_oracle(case, dtype, oracle):
    from lbfgsb_amd import problems as pb
    from lbfgsb_amd import solver
    name, p, eps, mi, tol32, exact = _cases()[case]
    print(f"case {name} {np.dtype(dtype).name}", flush=True)
    ref = oracle.Solver(p, dtype=np.float64, eps=eps, max_iter=mi).run()
    dp = solver.DeviceProblem.create(p, dtype)
    x, summ = solver.minimize(dp, eps=eps, max_iter=mi)
    xg = x.double().cpu().numpy()
    f_gpu, _ = pb.objective_numpy(p, xg)
    f_ref, _ = pb.objective_numpy(p, ref.arr("x"))
    lo = np.where((p.nbd == 1) | (p.nbd == 2), p.l, -np.inf)
    hi = np.where((p.nbd == 2) | (p.nbd == 3), p.u, np.inf)
    slack = 1e-12 if dtype == np.float64 else 1e-6
    assert np.all(xg >= lo - slack) and np.all(xg <= hi + slack)
    rel = rel_error(f_gpu, f_ref, p.kind)
    print(name, np.dtype(dtype).name, "gpu f", f_gpu, "oracle f", f_ref, "rel", rel,
          "iters", summ["iters"], ref.st.iter, "info", summ["info"], ref.st.info, flush=True)
    if dtype == np.float64:
        assert summ["info"] == ref.st.info
        if exact:
            assert summ["iters"] == ref.st.iter and rel < 1e-12
        else:
            assert abs(summ["iters"] - ref.st.iter) <= 2
            assert rel < solve_tolerance(eps, f_ref, p.kind, 1e-12)
    else:
        assert summ["info"] in (1, 2, 4, 5), summ
        assert rel < tol32


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_device_value_path_matches_callback(dtype):
    """The library's own objective as funcgrad (each trial's value read by
    the line search from device memory) against the same objective inside a
    foreign callback (the value returned on the host): the same solve, bit
    for bit."""
    from lbfgsb_amd import problems as pb
    from lbfgsb_amd import solver
    for p, eps in ((pb.rosenbrock(4096), 1e-7), (pb.grid_cos(64), 1e-7),
                   (pb.penta_qp(5000), 1e-6)):
        print(f"case kind {p.kind} n {p.n} {np.dtype(dtype).name}", flush=True)
        dp = solver.DeviceProblem.create(p, dtype)
        xa, sa = solver.minimize(dp, eps=eps, max_iter=2000)
        xb, sb = solver.minimize(dp, eps=eps, max_iter=2000, foreign_funcgrad=True)
        dp.close()
        assert sa == sb
        assert np.array_equal(xa.cpu().numpy(), xb.cpu().numpy())
