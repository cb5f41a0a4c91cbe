"""Failure, restart and exit paths of the driver on the GPU against the CPU
oracle (lbfgsbcpu.cpp:258-278, 297-341, 359-387, 394-460).

Each condition is forced from the host, so every kernel still runs to its
normal end:
  * a loaded oracle state with a modified memory (theta < 0, S'S negated),
    the same state on both sides;
  * evaluations replaced by chosen f and g values on both sides
    (tests/lockstep.py `feed`);
  * bad input, tolerances and callbacks through lbfgsb_minimize_*.

Each forcing state was checked on the oracle to take the path the test
names (the test asserts it again).  Comparisons are those of
tests/lockstep.py: decisions, counters and info exact, values within the
bulk precision's rtol.
"""
import dataclasses

import numpy as np
import pytest

from conftest import gpu_available
from lockstep import Lockstep, as_f64, load_state, states

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not gpu_available(), reason="needs a ROCm GPU")]

DTYPES = [np.float64, np.float32]


def _loaded(oracle, p, dtype, k, mod=None, spec=True):
    """Lockstep from the oracle's state at the top of iteration k, with
    `mod` applied to that state before both sides take it."""
    from lbfgsb_amd import solver
    (_, s), = states(oracle, p, dtype, (k,))
    if mod is not None:
        mod(s)
    dp = solver.DeviceProblem.create(p, dtype)
    eng = solver.Engine(dp)
    load_state(eng, s, p)
    ref = s.clone() if np.dtype(dtype) == np.float64 else as_f64(oracle, s, p)
    ls = Lockstep(oracle, p, dtype, eps=1e-10, max_iter=200, spec=spec, eng=eng, ref=ref,
                  dp=dp)
    ls.status = eng.run(solver.PROG_ITER_START)
    return ls


def _neg_theta(s):
    s.st.theta = -abs(s.st.theta)


@pytest.mark.parametrize("spec", [True, False], ids=["spec", "regular"])
@pytest.mark.parametrize("dtype", DTYPES)
def test_nondescent_with_empty_memory_ends_minus2(oracle, dtype, spec):
    """theta < 0 with col = 0: the Cauchy step is zero, g'd = 0 >= 0 (-4),
    and with no memory to discard the solve ends with info -2, x and f
    restored (lbfgsbcpu.cpp:359-375, 1307-1311)."""
    from lbfgsb_amd import problems as pb
    ls = _loaded(oracle, pb.penta_qp(4000, seed=102), dtype, 0, _neg_theta, spec)
    assert not ls.iteration()
    assert ls.rs.info == -2 and int(ls.es("info")) == -2
    assert int(ls.es("iter")) == ls.rs.iter == 1
    assert ls.es("f") == ls.rs.f  # f restored to fold
    ls.close()


@pytest.mark.parametrize("spec", [True, False], ids=["spec", "regular"])
@pytest.mark.parametrize("dtype", DTYPES)
def test_nondescent_with_memory_restarts(oracle, dtype, spec):
    """theta scaled by 10 on grid_cos 20x20 at iteration 5: the subspace
    backtrack gives alpha = 0, d = 0, g'd = 0 (-4) with col = 5, so the
    memory is discarded and the iteration starts again (lbfgsbcpu.cpp:
    376-386); then three more iterations in lockstep."""
    from lbfgsb_amd import problems as pb

    def scale(s):
        s.st.theta = s.st.theta * 10.0
    ls = _loaded(oracle, pb.grid_cos(20, seed=9), dtype, 5, scale, spec)
    for _ in range(4):
        assert ls.iteration()
    assert ls.counts["restarts"] == 1
    ls.close()


@pytest.mark.parametrize("dtype", DTYPES)
def test_formk_failure_restarts(oracle, dtype):
    """theta < 0 with a memory: formk's factorisation fails and the memory
    is discarded before the subspace step (lbfgsbcpu.cpp:313-325)."""
    from lbfgsb_amd import problems as pb
    ls = _loaded(oracle, pb.penta_qp(4000, seed=102), dtype, 3, _neg_theta)
    r0 = ls.rs.restarts
    # In fp32 the restarted search's first step (1) exceeds stpmx (1 - 5e-8),
    # dcsrch returns its error code, and the search resumes from the
    # previous search's saved state, as in the reference; the step it takes
    # raises f and the decrease test ends the solve on both sides.
    for _ in range(3):
        if not ls.iteration():
            break
    assert ls.rs.restarts == r0 + 1 == int(ls.es("restarts"))
    ls.close()


@pytest.mark.parametrize("dtype", DTYPES)
def test_formt_failure_resets(oracle, dtype):
    """S'S negated: the step is taken, but formt's Cholesky of
    T = theta S'S + L D^-1 L' fails after the update and the memory is
    discarded (lbfgsbcpu.cpp:446-459)."""
    from lbfgsb_amd import problems as pb

    def neg_ss(s):
        s.arr("ss")[:] *= -1.0
    ls = _loaded(oracle, pb.penta_qp(4000, seed=102), dtype, 3, neg_ss)
    r0 = ls.rs.restarts
    assert ls.iteration()
    assert ls.rs.restarts == r0 + 1 == int(ls.es("restarts"))
    assert ls.rs.col == 0 == int(ls.es("col"))
    for _ in range(2):
        assert ls.iteration()
    ls.close()


@pytest.mark.parametrize("spec", [True, False], ids=["spec", "regular"])
@pytest.mark.parametrize("dtype", DTYPES)
def test_twenty_backtracks(oracle, dtype, spec):
    """Every trial reports f = f_old + |g'd_0| stp with g = -g_0 (a line the
    search can only shrink): after 20 trials the reference leaves the loop
    and takes the 21st trial point unevaluated (lbfgsbcpu.cpp:352), here
    ending on eps_x."""
    from lbfgsb_amd import problems as pb
    ls = _loaded(oracle, pb.penta_qp(4000, seed=102), dtype, 2, None, spec)
    g0 = ls.eng.array("g").astype(np.float64)

    def feed(lk, f, x_t, g):
        return lk.rs.fold + abs(lk.rs.gdold) * lk.rs.stp, -g0
    ls.feed = feed
    assert not ls.iteration()
    assert ls.rs.iback == 20 and int(ls.es("iback")) == 20
    assert ls.rs.info == int(ls.es("info")) == 2
    ls.close()


@pytest.mark.parametrize("spec", [True, False], ids=["spec", "regular"])
@pytest.mark.parametrize("dtype", DTYPES)
def test_unit_step_after_leaving_z(oracle, dtype, spec):
    """Iteration 0 of an unboxed problem starts at stp = 1/|d| < 1 (stpmx =
    1).  Values with sufficient decrease and g'd = g'd_0 make the search
    extrapolate to stp = stpmx = 1 after two non-unit trials, so the trial
    pass forms the subspace point z again (the first trial overwrote it)
    and the trial point x = z (lbfgsbcpu.cpp:1329-1335).  Then up to three
    ordinary iterations."""
    from lbfgsb_amd import problems as pb
    ls = _loaded(oracle, pb.grid_cos(20, seed=9), dtype, 0, None, spec)
    g0 = ls.eng.array("g").astype(np.float64)
    steps = []

    def feed(lk, f, x_t, g):
        steps.append(lk.rs.stp)
        return lk.rs.fold + 0.5 * lk.rs.stp * lk.rs.gdold, g0
    ls.feed = feed
    assert ls.iteration()
    assert steps[0] < 1.0 and steps[-1] == 1.0 and len(steps) >= 3, steps
    ls.feed = None
    # the forced values leave an f_old the true f does not match, so the
    # decrease test may end the solve; lockstep checks both end alike
    for _ in range(3):
        if not ls.iteration():
            break
    ls.close()


@pytest.mark.parametrize("dtype", DTYPES)
def test_bad_input_minus1(oracle, dtype):
    """errclb (lbfgsbcpu.cpp:258-264, 871-905): nbd outside 0..3, or l > u
    with nbd = 2, gives info -1 and leaves x untouched."""
    from lbfgsb_amd import problems as pb
    from lbfgsb_amd import solver
    p = pb.penta_qp(3000, seed=11)
    nbd = p.nbd.copy()
    nbd[17] = 4
    lo, hi = p.l.copy(), p.u.copy()
    lo[123], hi[123] = 0.5, -0.5
    for q in (dataclasses.replace(p, nbd=nbd), dataclasses.replace(p, l=lo, u=hi)):
        ref = oracle.minimize(q, dtype=np.float64, eps=1e-6)
        x, summ = solver.minimize(solver.DeviceProblem.create(q, dtype), eps=1e-6)
        assert ref["info"] == -1 and summ["info"] == -1
        assert np.array_equal(x.cpu().numpy(), q.cast(dtype).x0)


@pytest.mark.parametrize("dtype", DTYPES)
def test_initial_projected_gradient_stop(oracle, dtype):
    """sbgnrm <= eps_g at the projected x0 (lbfgsbcpu.cpp:271-278): info 4,
    no iteration, x = the projection of x0 (exact)."""
    from lbfgsb_amd import problems as pb
    from lbfgsb_amd import solver
    p = pb.rosenbrock(4000, seed=12)  # x0 ~ U(-1.2, 1.2) leaves [-0.5, 1.5] on even indices
    ls = Lockstep(oracle, p, dtype, eps=1e-8, eps_g=1e4)
    assert not ls.start()
    assert int(ls.es("prjctd")) == 1 and ls.rs.info == 4
    ls.close()
    ref = oracle.minimize(p, dtype=dtype, eps=1e-8, eps_g=1e4)
    x, summ = solver.minimize(solver.DeviceProblem.create(p, dtype), eps=1e-8, eps_g=1e4)
    assert summ["info"] == ref["info"] == 4 and summ["iters"] == ref["iters"] == 0
    assert np.array_equal(x.cpu().numpy(), ref["x"])


@pytest.mark.parametrize("dtype", DTYPES)
def test_customized_stop_and_max_iteration(oracle, dtype):
    """A customised stopping test that returns non-zero at iteration 3 ends
    with info 0 (lbfgsbcpu.cpp:423-429); max_iteration = 3 ends with info 5
    after iteration 4 (:418-422).  Both through lbfgsb_minimize, against
    the oracle's solve with the same stop."""
    from lbfgsb_amd import problems as pb
    from lbfgsb_amd import solver
    from lockstep import close
    p = pb.grid_cos(40, seed=13)
    dp = solver.DeviceProblem.create(p, dtype)
    rtol = 1e-9 if dtype == np.float64 else 2e-4
    ref = oracle.minimize(p, dtype=np.float64, eps=1e-10, stop_at_iter=3)
    x, summ = solver.minimize(dp, eps=1e-10, stop=lambda s: s.num_iteration >= 3)
    assert summ["info"] == ref["info"] == 0 and summ["iters"] == ref["iters"] == 3
    close(x.double().cpu().numpy(), ref["x"], 100 * rtol, "x at the custom stop")
    ref = oracle.minimize(p, dtype=np.float64, eps=1e-10, max_iter=3)
    x, summ = solver.minimize(dp, eps=1e-10, max_iter=3)
    assert summ["info"] == ref["info"] == 5 and summ["iters"] == ref["iters"] == 4
    close(x.double().cpu().numpy(), ref["x"], 100 * rtol, "x at max_iteration")
    # an after-iteration callback sees every accepted iterate and changes nothing
    seen = []
    x2, summ2 = solver.minimize(dp, eps=1e-10, max_iter=3,
                                after_iteration=lambda s: seen.append(s.num_iteration))
    assert seen == [1, 2, 3, 4] and summ2["iters"] == 4
    assert np.array_equal(x2.cpu().numpy(), x.cpu().numpy())
