"""The whole-chain program after an evaluation runs pass A on the trial point
before the line-search step, with the new S/Y pair in a spare ring slot.  It
must take the same solve as the regular program (g'd pass, line-search step,
then pass A on the accepted point), which a partial run naming the program's
first step selects: the same iterations, ring positions and decisions, and
the same values up to the last bits of g'd, which the two forms sum over
different block partitions (both compensated).  Forcing the memory update to
be skipped at some iterations takes the speculative form through its redo
status (the host then runs the after-acceptance program)."""
import numpy as np
import pytest

from conftest import gpu_available

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not gpu_available(), reason="needs a ROCm GPU")]


def _drive(eng, regular, skip_iters, eps=1e-10, max_iter=40, max_evals=400):
    """Engine.solve's host loop; `regular` runs program 1 from its first
    step (the regular form).  Before each program-1 run at an iteration in
    skip_iters, gdold is made huge so that accept_step skips the memory
    update (dr <= eps * ddum, lbfgsbcpu.cpp:433-436)."""
    from lbfgsb_amd import solver as S
    st = eng.setup(eps, max_iter)
    prog, f = S.PROG_ITER_START, 0.0
    redo = 0
    while st not in (S.ST_DONE, -1) and eng.evals < max_evals:
        if prog == S.PROG_AFTER_EVAL:
            if int(eng.scalar("iter")) in skip_iters:
                eng.set_scalar("gdold", 1e30)
            st = eng.run(prog, S.K_GD if regular else -1, S.K_END, f)
        else:
            st = eng.run(prog, f=f)
        if st == S.ST_NEED_EVAL:
            f = eng.evaluate(1)
            prog = S.PROG_AFTER_EVAL
        elif st == S.ST_RESTART:
            prog = S.PROG_ITER_START
        elif st in (S.ST_ACCEPTED, S.ST_REDO):
            redo += st == S.ST_REDO
            prog = S.PROG_AFTER_ACCEPT
        elif st == S.ST_CHECKED:
            prog = S.PROG_AFTER_CHECK
    assert st == S.ST_DONE, st
    return redo


def _cases():
    from lbfgsb_amd import problems as pb
    return [("rosenbrock_16384", pb.rosenbrock(16384, seed=3)),
            ("penta_qp_6000", pb.penta_qp(6000, seed=5)),
            ("grid_cos_48x40", pb.grid_cos(48, 40, seed=4))]


@pytest.mark.parametrize("skip", [(), (2, 5, 6, 11)])
@pytest.mark.parametrize("case", range(3))
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_spec_program_matches_regular(case, dtype, skip):
    from lbfgsb_amd import solver
    name, p = _cases()[case]
    print(f"case {name} {np.dtype(dtype).name} skip={skip}", flush=True)
    dp = solver.DeviceProblem.create(p, dtype)
    out = {}
    for regular in (False, True):
        eng = solver.Engine(dp)
        redo = _drive(eng, regular, set(skip))
        out[regular] = dict(
            redo=redo,
            x=eng.array("x"), g=eng.array("g"),
            ws=eng.array("ws"), wy=eng.array("wy"), sy=eng.array("sy"),
            ss=eng.array("ss"), wt=eng.array("wt"),
            scal={k: eng.scalar(k) for k in ("iter", "nfgv", "nskip", "col", "head", "itail",
                                             "iupdat", "info", "f", "theta", "sbgnrm")})
        eng.close()
    dp.close()
    spec, reg = out[False], out[True]
    assert reg["redo"] == 0
    if skip:
        # every forced skip that ended an iteration came back as a redo
        assert spec["redo"] >= 1 and spec["scal"]["nskip"] >= 1
    for k in ("iter", "nfgv", "nskip", "col", "head", "itail", "iupdat", "info"):
        assert spec["scal"][k] == reg["scal"][k], (k, spec["scal"], reg["scal"])
    # tolerance: a last-bit difference of g'd moves the FP64 small state by
    # a few ulps and the bulk vectors by at most a few of their own ulps
    bulk_rtol = 1e-12 if dtype == np.float64 else 4e-6
    for k in ("f", "theta", "sbgnrm"):
        a, b = spec["scal"][k], reg["scal"][k]
        assert abs(a - b) <= 1e-10 * max(abs(a), abs(b), 1e-30), (k, a, b)
    for k in ("x", "g", "ws", "wy"):
        a, b = spec[k].astype(np.float64), reg[k].astype(np.float64)
        scale = max(np.max(np.abs(b)), 1e-30)
        assert np.max(np.abs(a - b)) <= bulk_rtol * scale, k
    for k in ("sy", "ss", "wt"):
        a, b = spec[k], reg[k]
        scale = max(np.max(np.abs(b)), 1e-30)
        assert np.max(np.abs(a - b)) <= 1e-9 * scale, k
