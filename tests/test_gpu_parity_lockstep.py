"""GPU parity over whole iterations: the engine and the CPU oracle run in
lockstep (tests/lockstep.py), from the engine's own set-up to the end of the
solve, with the GPU's f and g handed to the oracle at every evaluation.

At every iteration the comparison covers the Cauchy point and free sets
(exact), the subspace step, the first line-search call (d'd, stpmx, stp),
every trial (stp, iback, ifun, nfgv exact; the trial point), the line
search's end, the stopping tests (info, iter exact; sbgnrm bit-exact), the
update-or-skip decision, the ring (col, head, itail), theta, S'Y, S'S, wt
and the new S/Y columns.  Both the speculative after-evaluation program (the
default chain of lbfgsb_minimize) and the regular one are driven.

Cases:
  * the four generators at small n, to the end of the solve with their
    workload tolerances (equal iteration count and info);
  * the reference example's problem (dsscfg, unconstrained: the shortcut of
    lbfgsbcpu.cpp:281-284 from the second iteration on);
  * the plain fp64 workload at its full size (256 x 256 grid quadratic,
    eps 1e-8, 1000 iterations), to the end;
  * a pentadiagonal QP with n = 2^18, whose first Cauchy search sorts more
    than 8192 candidates in one round (at least three merge passes).
"""
import numpy as np
import pytest

from conftest import gpu_available

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not gpu_available(), reason="needs a ROCm GPU")]


def _case(name):
    from lbfgsb_amd import problems as pb
    return {
        "grid_cos": (pb.grid_cos(50, 48, seed=101), 1e-7, 2000),
        "penta": (pb.penta_qp(4000, seed=102), 1e-6, 1000),
        "rosen": (pb.rosenbrock(4000, seed=103), 1e-7, 1000),
        "grid_quad": (pb.grid_quadratic(48, 50, seed=104), 1e-8, 1000),
        "dsscfg": (pb.dsscfg(40), 1e-8, 1000),
    }[name]


def _report(ls, name):
    print(f"  {name}: iters {ls.iters} info {ls.rs.info} trials {ls.counts['trials']} "
          f"restarts {ls.counts['restarts']} skips {ls.counts['skips']} "
          f"sorted rounds {ls.counts['rounds']} max candidates {ls.counts['max_cand']}",
          flush=True)


@pytest.mark.parametrize("spec", [True, False], ids=["spec", "regular"])
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("name", ["grid_cos", "penta", "rosen", "grid_quad", "dsscfg"])
def test_lockstep_to_the_end(oracle, name, dtype, spec):
    from lockstep import Lockstep
    p, eps, mi = _case(name)
    print(f"case {name} {np.dtype(dtype).name} spec={spec}", flush=True)
    ls = Lockstep(oracle, p, dtype, eps=eps, max_iter=mi, spec=spec)
    more = ls.run(max_iters=mi + 1)
    _report(ls, name)
    assert not more, "solve did not end"
    assert ls.rs.info in (1, 2, 4, 5)
    assert int(ls.es("iter")) == ls.rs.iter
    if name == "dsscfg":
        assert int(ls.es("shortcut")) == 1  # the unconstrained shortcut ran
    ls.close()


def test_lockstep_grid_quadratic_256(oracle):
    """The plain fp64 workload at its full size (n = 65,536), to the end."""
    from lbfgsb_amd import problems as pb
    from lockstep import Lockstep
    print("case grid_quadratic 256x256 fp64", flush=True)
    ls = Lockstep(oracle, pb.grid_quadratic(256), np.float64, eps=1e-8, max_iter=1000)
    more = ls.run(max_iters=1001)
    _report(ls, "grid_quadratic_256")
    assert not more and ls.rs.info in (1, 2, 4, 5)
    ls.close()


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_lockstep_large_cauchy_round(oracle, dtype):
    """n = 2^18: the first Cauchy search passes ~1.5e5 breakpoints, sorted
    by tiles and merge-path passes on the GPU."""
    from lbfgsb_amd import problems as pb
    from lockstep import Lockstep
    print(f"case penta 2^18 {np.dtype(dtype).name}", flush=True)
    ls = Lockstep(oracle, pb.penta_qp(1 << 18, seed=105), dtype, eps=1e-10, max_iter=200)
    ls.run(max_iters=4)
    _report(ls, "penta_262144")
    assert ls.iters == 4
    assert ls.counts["rounds"] >= 1 and ls.counts["max_cand"] > 4 * 2048
    ls.close()
