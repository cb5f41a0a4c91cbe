"""The allocation-free entry point (lbfgsb_minimize_noalloc_*) against
lbfgsb_minimize_* through the C ABI: the same kernels in the same order on
buffers carved from the per-GPU scratch region, so the results are
bit-identical (x, iterations, info, residuals).  Two calls in a row, the
second smaller, show that what the region keeps from one call does not
reach the next; a size beyond the region is refused with x untouched."""
import numpy as np
import pytest

from conftest import gpu_available

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not gpu_available(), reason="needs a ROCm GPU")]


def _same(dp, eps, mi):
    from lbfgsb_amd import solver
    xa, sa = solver.minimize(dp, eps=eps, max_iter=mi)
    xb, sb = solver.minimize(dp, eps=eps, max_iter=mi, noalloc=True)
    print(f"  {dp.problem.name}: iters {sa['iters']}/{sb['iters']} info {sa['info']}/{sb['info']}",
          flush=True)
    assert sb["rc"] == sa["info"] == sb["info"] and sa["info"] in (1, 2, 4, 5)
    assert sa["iters"] == sb["iters"] > 0
    for k in ("residual_g", "residual_f", "residual_x"):
        assert sa[k] == sb[k], k
    assert np.array_equal(xa.cpu().numpy(), xb.cpu().numpy())


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_noalloc_matches_minimize_twice(dtype):
    from lbfgsb_amd import problems as pb
    from lbfgsb_amd import solver
    print(f"case noalloc {np.dtype(dtype).name}", flush=True)
    big = pb.penta_qp(200000, seed=71)       # most of the fp64 region
    small = pb.grid_cos(64, 48, seed=72)     # then a smaller solve in the same region
    for p, eps, mi in ((big, 1e-6, 1000), (small, 1e-7, 2000)):
        _same(solver.DeviceProblem.create(p, dtype), eps, mi)


def test_noalloc_refuses_beyond_the_region():
    from lbfgsb_amd import problems as pb
    from lbfgsb_amd import solver
    print("case noalloc too large", flush=True)
    p = pb.rosenbrock(400000)                # fp64 region holds 291,584 variables
    dp = solver.DeviceProblem.create(p, np.float64)
    x, s = solver.minimize(dp, eps=1e-6, max_iter=10, noalloc=True)
    assert s["rc"] == -3 and s["info"] == -3
    assert np.array_equal(x.cpu().numpy(), p.x0)
