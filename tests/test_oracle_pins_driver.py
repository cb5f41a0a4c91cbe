"""Pins of the oracle's driver-level parts against their definitions:
active (lbfgsbcpu.cpp:467-515), the first lnsrlb call's largest feasible
step and first step (:1263-1304), and the order of the stopping tests with
their info codes (:394-429).

Nothing here restates the oracle's loops: the projection is numpy's clip
onto the box each bound type describes, the largest feasible step is found
by bisection on feasibility of x + t d, and the stopping order is read off
solves whose settings make several tests true at the same iteration.
"""
import dataclasses

import numpy as np
import pytest

from lbfgsb_amd import problems as pb

RNG = np.random.default_rng(7)


def _box_problem(n=400, seed=3):
    """A pentadiagonal QP with every bound type, some l = u, and a start
    point partly outside the box."""
    p = pb.penta_qp(n, seed=seed)
    nbd = RNG.integers(0, 4, n).astype(np.int32)
    lo = RNG.uniform(-1.0, 0.0, n)
    hi = lo + RNG.uniform(0.0, 1.0, n)
    fixed = RNG.random(n) < 0.05
    nbd[fixed] = 2
    hi[fixed] = lo[fixed]                      # l = u
    x0 = RNG.uniform(-1.5, 1.5, n)
    return dataclasses.replace(p, nbd=nbd, l=lo, u=hi, x0=x0)


def test_active_projects_and_classifies(oracle):
    p = _box_problem()
    s = oracle.Solver(p, eps=1e-30)
    st = s.st
    lo = np.where((p.nbd == 1) | (p.nbd == 2), p.l, -np.inf)
    hi = np.where((p.nbd == 2) | (p.nbd == 3), p.u, np.inf)
    proj = np.clip(p.x0, lo, hi)
    assert np.array_equal(s.arr("x"), proj)  # the start point is projected, exactly
    iw = s.arr("iwhere")
    want = np.where(p.nbd == 0, -1, np.where((p.nbd == 2) & (p.u - p.l <= 0), 3, 0))
    assert np.array_equal(iw, want)
    assert np.count_nonzero(want == 3) > 0 and np.count_nonzero(want == -1) > 0
    assert st.prjctd == int(np.any(proj != p.x0))
    assert st.cnstnd == int(np.any(p.nbd != 0))
    assert st.boxed == int(np.all(p.nbd == 2))
    # the flags in the other cases
    q = dataclasses.replace(p, nbd=np.full(p.n, 2, np.int32), l=p.x0 - 1, u=p.x0 + 1)
    s = oracle.Solver(q, eps=1e-30)
    assert (s.st.prjctd, s.st.cnstnd, s.st.boxed) == (0, 1, 1)
    q = dataclasses.replace(p, nbd=np.zeros(p.n, np.int32))
    s = oracle.Solver(q, eps=1e-30)
    assert (s.st.prjctd, s.st.cnstnd, s.st.boxed) == (0, 0, 0)
    assert np.all(s.arr("iwhere") == -1)


def _feasible(x, d, t, p, tol):
    y = x + t * d
    okl = np.where((p.nbd == 1) | (p.nbd == 2), y >= p.l - tol, True)
    oku = np.where((p.nbd == 2) | (p.nbd == 3), y <= p.u + tol, True)
    return bool(np.all(okl & oku))


def _largest_step(x, d, p, big=1e10):
    """Largest t in [0, big] with x + t d inside the box, by bisection on
    feasibility (relative resolution 1e-13)."""
    scale = np.max(np.abs(x)) + 1.0
    tol = 1e-13 * scale
    if _feasible(x, d, big, p, tol):
        return big
    lo_t, hi_t = 0.0, big
    if not _feasible(x, d, 0.0, p, tol):
        return 0.0
    for _ in range(400):
        mid = 0.5 * (lo_t + hi_t)
        if _feasible(x, d, mid, p, tol):
            lo_t = mid
        else:
            hi_t = mid
        if hi_t - lo_t <= 1e-13 * hi_t:
            break
    return lo_t


def _first_lnsrlb(oracle, p, x, d, g, it, cnstnd, boxed):
    s = oracle.Solver(p, eps=1e-30)
    st = s.st
    s.arr("x")[:] = x
    s.arr("d")[:] = d
    s.arr("z")[:] = x + d
    s.arr("g")[:] = g
    st.iter, st.cnstnd, st.boxed, st.task = it, cnstnd, boxed, 0
    s.fl.lnsrlb_step(s.s)
    return s


@pytest.mark.parametrize("trial", range(6))
def test_lnsrlb_largest_feasible_step(oracle, trial):
    """After the first iteration of a constrained problem, stpmx is the
    largest step along d that keeps every bounded variable in its box
    (capped at big = 1e10); 0 when a variable sits on a bound d leaves."""
    p = _box_problem(seed=10 + trial)
    lo = np.where((p.nbd == 1) | (p.nbd == 2), p.l, -np.inf)
    hi = np.where((p.nbd == 2) | (p.nbd == 3), p.u, np.inf)
    # strictly inside the box, so that the bounds d reaches decide
    x = np.clip(RNG.uniform(-1.0, 1.0, p.n), lo + 1e-3, hi - 1e-3)
    d = RNG.standard_normal(p.n) * 0.1
    fixed = p.u - p.l <= 0
    x[fixed] = p.l[fixed]
    d[fixed] = 0.0  # a fixed variable does not move
    if trial == 3:  # a variable on its lower bound, d pointing out of the box
        k = int(np.where(p.nbd == 1)[0][0])
        x[k], d[k] = p.l[k], -0.5
    if trial == 4:  # no bound in the direction of d: big
        d = np.where(p.nbd == 0, d, 0.0)
    if trial == 5:  # fixed variables (l = u) with d = 0 there, others free
        d = np.where(p.u - p.l <= 0, 0.0, d)
        d = np.where(p.nbd == 0, d, 0.0)
        d[np.where(p.nbd == 0)[0][:3]] = 1.0
    g = -d  # a descent direction
    s = _first_lnsrlb(oracle, p, x, d, g, it=2, cnstnd=1, boxed=0)
    want = _largest_step(x, d, p)
    if want < 1e-11:
        want = 0.0  # within the bisection's feasibility tolerance of 0
    got = s.st.stpmx
    if want in (0.0, 1e10):
        assert got == want, (got, want)
    else:
        # the bisection resolves feasibility to 1e-13 of |x|, i.e. ~1e-12 in t
        assert abs(got - want) <= 1e-9 * want, (got, want)
    assert s.st.stp == 1.0  # after the first iteration the search starts at 1
    assert np.isclose(s.st.dtd, d @ d, rtol=1e-14)


@pytest.mark.parametrize("case", ["unboxed", "boxed", "unconstrained", "later"])
@pytest.mark.parametrize("dnorm", [0.25, 40.0])
def test_lnsrlb_first_step(oracle, case, dnorm):
    """The first step: min(1/|d|, stpmx) in iteration 0 of a problem that
    is not boxed, with stpmx = 1 when it is constrained and 1e10 when not;
    1 in iteration 0 of a boxed problem and in every later iteration.  The
    trial point is z for a unit step and x + stp d otherwise."""
    p = _box_problem(seed=31)
    x = np.clip(RNG.uniform(-1, 1, p.n), p.l - 10, p.u + 10)
    d = RNG.standard_normal(p.n)
    d *= dnorm / np.linalg.norm(d)
    it = 1 if case == "later" else 0
    cnstnd = 0 if case == "unconstrained" else 1
    boxed = 1 if case == "boxed" else 0
    s = _first_lnsrlb(oracle, p, x, d, -d, it, cnstnd, boxed)
    if case in ("unboxed", "unconstrained"):
        stpmx = 1.0 if cnstnd else 1e10
        want = min(1.0 / np.linalg.norm(d), stpmx)
        assert s.st.stpmx == stpmx
        assert np.isclose(s.st.stp, want, rtol=1e-14)
    else:
        assert s.st.stp == 1.0
    xt = s.arr("x")
    if s.st.stp == 1.0:
        assert np.array_equal(xt, x + d)  # z
    else:
        assert np.allclose(xt, x + s.st.stp * d, rtol=1e-15, atol=1e-15)
    assert s.st.task == 1 and s.st.iback == 0 and s.st.ifun == 1


def _end(oracle, p, **kw):
    r = oracle.minimize(p, eps=0.0, **kw)
    return r["info"], r["iters"]


def test_stopping_order_and_codes(oracle):
    """Several stopping tests true at the same iteration: the first in the
    reference's order decides (projected gradient 4, step 2, decrease 1,
    iteration count 5, customised test 0)."""
    p = pb.grid_cos(30, seed=5)
    ref = oracle.Solver(p, eps=1e-30).run(3)
    sb0 = oracle.Solver(p, eps=1e-30).st.sbgnrm
    sb1 = ref.records[0].sbgnrm
    assert sb1 < sb0
    eg = 0.5 * (sb0 + sb1)  # passes after iteration 1, not at x0
    big = 1e30
    assert _end(oracle, p, eps_g=eg, eps_x=big, eps_f=big) == (4, 1)
    assert _end(oracle, p, eps_g=0.0, eps_x=big, eps_f=big) == (2, 1)
    assert _end(oracle, p, eps_g=0.0, eps_x=0.0, eps_f=big) == (1, 1)
    assert _end(oracle, p, eps_g=0.0, eps_x=0.0, eps_f=big, max_iter=1, stop_at_iter=1) == (1, 1)
    assert _end(oracle, p, eps_g=0.0, eps_x=0.0, eps_f=0.0, max_iter=1, stop_at_iter=2) == (5, 2)
    assert _end(oracle, p, eps_g=0.0, eps_x=0.0, eps_f=0.0, max_iter=1, stop_at_iter=1) == (0, 1)
    assert _end(oracle, p, eps_g=0.0, eps_x=0.0, eps_f=0.0, max_iter=3) == (5, 4)
    # at x0: only the projected-gradient test runs (lbfgsbcpu.cpp:271-278)
    assert _end(oracle, p, eps_g=2 * sb0, eps_x=big, eps_f=big) == (4, 0)
