"""Lockstep driver for GPU parity: the HIP engine and the CPU oracle take the
same iterations side by side, and every decision and value is compared at
the point where the reference's driver produces it (lbfgsbcpu.cpp:279-462).

Only the function evaluations are shared: at each trial point the GPU's
built-in objective gives f and g, and the oracle is handed the GPU's trial
point, f and g (after its own trial point has been compared with the GPU's).
The oracle then takes every other step by itself -- active, projgr, cauchy,
freev, formk, cmprlb, subsm, lnsrlb/dcsrch, the stopping tests, matupd,
formt -- so a difference in any of them shows at the step that made it.

Between the subspace step and the line search the oracle takes the GPU's z
and d = z - x, once they have been compared: d is a small difference of
large numbers late in a solve, and the GPU's float or differently summed z
would otherwise move every later line-search value by far more than either
side's own error.

Tolerances: integer and set outputs (iwhere, index, indx2, nfree, nint,
nenter, ileave, wrk, iback, ifun, nfgv, iter, col, head, itail, updatd,
nskip, restarts, info) are exact.  The projected-gradient norm is a max, so
it is compared bit for bit against the oracle's projgr evaluated in the bulk
precision on the GPU's own x and g.  Floating-point values use rtol relative
to the largest entry: 1e-9 for fp64.  The fp32 engine keeps FP32 n-vectors
but FP64 scalar and small-dense logic, so its reference is the FP64 oracle
on the float-rounded problem, and values are compared with rtol 2e-4 (the
float rounding of the trial points and S/Y pairs).
"""
from __future__ import annotations

import numpy as np

from lbfgsb_amd import solver as S

TOL = {np.dtype(np.float64): 1e-9, np.dtype(np.float32): 2e-4}


def close(a, b, rtol, what):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    if a.size == 0:
        return 0.0
    scale = max(float(np.max(np.abs(b))), 1e-300)
    err = float(np.max(np.abs(a - b))) / scale
    assert err <= rtol, f"{what}: max rel err {err:.3e} > {rtol}"
    return err


def scalar_close(a, b, rtol, what):
    a, b = float(a), float(b)
    assert abs(a - b) <= rtol * max(abs(b), 1e-300), f"{what}: {a!r} vs {b!r}"


def load_state(engine, s, p):
    """Copy the oracle state (top of an iteration) into the engine."""
    st = s.st
    n, m = st.n, st.m
    dt = engine.np_dtype
    engine.set_array("x", s.arr("x").astype(dt))
    engine.set_array("g", s.arr("g").astype(dt))
    engine.set_array("ws", s.arr("ws").astype(dt))
    engine.set_array("wy", s.arr("wy").astype(dt))
    for name in ("sy", "ss", "wt"):
        engine.set_array(name, s.arr(name).astype(np.float64))
    engine.set_array("wn", s.arr("wn").astype(np.float64))
    engine.set_array("snd", s.arr("snd").astype(np.float64))
    c = np.zeros(16)
    wa = np.ctypeslib.as_array(st.wa)
    c[: 2 * st.col] = wa[2 * m: 2 * m + 2 * st.col]
    engine.set_array("c", c)
    # the line search's saved state: a search whose first step exceeds
    # stpmx resumes from the previous search's (lbfgsbcpu.cpp:1574-1608)
    engine.set_array("dsave", np.array([float(v) for v in st.dsave]))
    engine.set_array("isave", np.array([float(v) for v in st.isave]))
    iw = s.arr("iwhere").astype(np.int64)
    free_last = np.zeros(n, bool)
    idx = s.arr("index")
    free_last[idx[: st.nfree] - 1] = True
    bits = ((iw + 4) | (free_last.astype(np.int64) << 4) | (free_last.astype(np.int64) << 5)
            | (p.nbd.astype(np.int64) << 6))  # the class byte carries nbd
    engine.set_array("iwbits", bits.astype(np.uint8))
    vals = dict(col=st.col, head=st.head - 1, itail=max(st.itail - 1, 0), iupdat=st.iupdat,
                iter=st.iter, updatd=st.updatd, theta=st.theta, f=st.f, cnstnd=st.cnstnd,
                boxed=st.boxed, nfree=st.nfree, sbgnrm=st.sbgnrm, nint=st.nint, nfgv=st.nfgv,
                nskip=st.nskip, restarts=st.restarts, phase=0,
                status=0, halt=0, eps_g=1e-10, eps_f=1e-10, eps_x=1e-10, max_iteration=200,
                has_after_iteration=0, has_custom_stop=0, upd_done=0)
    for k, v in vals.items():
        engine.set_scalar(k, v)


_SCALARS_I = ["col", "head", "itail", "iupdat", "iter", "nfgv", "nint", "nintol", "nskip",
              "nfree", "nact", "ileave", "nenter", "iword", "ifun", "iback", "info", "task",
              "csave", "internalinfo", "prjctd", "cnstnd", "boxed", "updatd", "wrk", "k_err",
              "restarts"]
_SCALARS_R = ["f", "fold", "theta", "dr", "rr", "dnrm", "xstep", "sbgnrm", "ddum", "dtd", "gd",
              "gdold", "stp", "stpmx", "tf"]


def as_f64(oracle, s32, p):
    """The float32 oracle state widened to an FP64 oracle state: the
    reference for the mixed-precision engine (FP32 data, FP64 scalar and
    small-dense logic)."""
    p64 = p.cast(np.float32).cast(np.float64)
    s = oracle.Solver(p64, dtype=np.float64, eps=1e-10, max_iter=200)
    a, b = s32.st, s.st
    for k in _SCALARS_I:
        setattr(b, k, getattr(a, k))
    for k in _SCALARS_R:
        setattr(b, k, float(getattr(a, k)))
    for i in range(2):
        b.isave[i] = a.isave[i]
    for i in range(13):
        b.dsave[i] = float(a.dsave[i])
    for i in range(64):
        b.wa[i] = float(a.wa[i])
    for name in ("x", "g", "l", "u", "xold", "xdiff", "z", "r", "d", "t", "ws", "wy", "sy",
                 "ss", "wt", "wn", "snd"):
        s.arr(name)[:] = s32.arr(name).astype(np.float64)
    for name in ("nbd", "index", "iwhere", "indx2"):
        s.arr(name)[:] = s32.arr(name)
    return s


def states(oracle, p, dtype, iters):
    s = oracle.Solver(p, dtype=dtype, eps=1e-10, max_iter=200)
    out, k = [], 0
    while s.more and k <= max(iters):
        if k in iters:
            out.append((k, s.clone()))
        s.step()
        k += 1
    return out


class Lockstep:
    """One engine and one oracle state driven together.

    Build with `start()` (the engine's own set-up against the oracle's) or
    hand in an engine and an oracle state that are already matched (a
    loaded state), then call `iteration()` repeatedly."""

    def __init__(self, oracle, p, dtype, eps=1e-10, max_iter=200, spec=True, m=8,
                 eps_g=None, eps_f=None, eps_x=None, eng=None, ref=None, dp=None):
        self.oracle = oracle
        self.p = p
        self.dtype = np.dtype(dtype)
        self.f64 = self.dtype == np.float64
        self.rtol = TOL[self.dtype]
        self.spec = spec
        self.m = m
        self.eps = dict(eps_g=eps if eps_g is None else eps_g,
                        eps_f=eps if eps_f is None else eps_f,
                        eps_x=eps if eps_x is None else eps_x)
        self.max_iter = max_iter
        # projgr in the bulk precision, for the bit-exact sbgnrm check
        self.pj = oracle.Solver(p, dtype=self.dtype, m=m, eps=1.0, max_iter=1)
        if ref is None:
            rp = p if self.f64 else p.cast(np.float32).cast(np.float64)
            ref = oracle.Solver(rp, dtype=np.float64, m=m, max_iter=max_iter, **self.eps)
        self.ref = ref
        self.dp = dp if dp is not None else S.DeviceProblem.create(p, self.dtype.type)
        self.eng = eng if eng is not None else S.Engine(self.dp, m=m)
        self.status = None
        self.iters = 0
        self.counts = dict(trials=0, restarts=0, skips=0, rounds=0, max_cand=0)
        # feed(lockstep, f, x_trial, g) -> (f, g or None): replaces the value
        # (and gradient) of an evaluation on both sides
        self.feed = None

    # ---------------------------------------------------------------- utils
    @property
    def rs(self):
        return self.ref.st

    def es(self, name):
        return self.eng.scalar(name)

    def bulk_sbgnrm(self, x, g):
        st = self.pj.st
        self.pj.arr("x")[:] = x
        self.pj.arr("g")[:] = g
        self.pj.fl.projgr(self.pj.s)
        return float(st.sbgnrm)

    def run_after_eval(self, f):
        e = self.eng
        if self.spec:
            st = e.run(S.PROG_AFTER_EVAL, f=f)
            if st == S.ST_REDO:
                st = e.run(S.PROG_AFTER_ACCEPT)
        else:
            st = e.run(S.PROG_AFTER_EVAL, S.K_GD, S.K_END, f=f)
        return st

    def note_rounds(self):
        self.counts["max_cand"] = int(self.es("max_sorted_round"))
        self.counts["rounds"] = int(self.es("gcp_rounds_run"))

    # --------------------------------------------------------------- set-up
    def start(self):
        """errclb + active + the first evaluation + projgr on both sides
        (lbfgsbcpu.cpp:258-278).  Returns False if the solve ended there."""
        e, ref, rs = self.eng, self.ref, self.rs
        st = e.setup(max_iter=self.max_iter, **self.eps)
        if st == S.ST_DONE and int(self.es("info")) == -1:
            # errclb: bad input, x untouched on both sides
            assert rs.info == -1, rs.info
            assert np.array_equal(e.array("x").astype(np.float64), ref.arr("x"))
            return False
        x_e = e.array("x")
        # active: projection of x0 and the initial iwhere, exact
        assert np.array_equal(x_e.astype(np.float64), ref.arr("x")), "active: projected x0"
        assert np.array_equal(e.array("iwhere"), ref.arr("iwhere")), "active: iwhere"
        for k in ("prjctd", "cnstnd", "boxed"):
            assert int(self.es(k)) == int(getattr(rs, k)), k
        # the shared evaluation at x0, then the oracle's projgr on it
        g_e = e.array("g")
        ref.arr("g")[:] = g_e
        rs.f = e.f
        ref.fl.projgr(ref.s)
        rs.residual_g = rs.sbgnrm
        assert self.es("sbgnrm") == self.bulk_sbgnrm(x_e, g_e), "projgr at x0 (bit-exact)"
        ended = rs.sbgnrm <= self.eps["eps_g"]
        if ended:
            rs.info = 4
            assert st == S.ST_DONE and int(self.es("info")) == 4, (st, self.es("info"))
            return False
        assert st != S.ST_DONE, "engine stopped at x0, oracle did not"
        self.status = e.run(S.PROG_ITER_START)
        return True

    # ------------------------------------------------------------ iteration
    def head(self):
        """The oracle's Cauchy point .. subspace step, compared with the
        GPU's chain (which already ran through its first trial point)."""
        e, ref, rs, rtol = self.eng, self.ref, self.rs, self.rtol
        r0 = rs.restarts
        ref.fl.iter_head(ref.s)
        # memory resets inside the head (cauchy / formk / subsm failures):
        # the GPU chain ended in a restart for each, and starts again
        for _ in range(rs.restarts - r0):
            assert self.status == S.ST_RESTART, ("head restart", self.status)
            self.status = e.run(S.PROG_ITER_START)
        self.note_rounds()
        shortcut = (not rs.cnstnd) and rs.col > 0
        if not shortcut:
            iw_g, iw_r = e.array("iwhere"), ref.arr("iwhere")
            bad = np.where(iw_g != iw_r)[0]
            assert bad.size == 0, f"iwhere: {bad.size} mismatches, first {bad[:8]}"
            assert int(self.es("nint")) == rs.nint, ("nint", self.es("nint"), rs.nint)
            assert int(self.es("nfree")) == rs.nfree, "nfree"
            assert int(self.es("wrk")) == int(rs.wrk), "wrk"
            assert np.array_equal(e.array("index"), ref.arr("index")), "index"
            if rs.iter > 0 and rs.cnstnd:
                assert int(self.es("nenter")) == rs.nenter, "nenter"
                assert int(self.es("ileave")) == rs.ileave, "ileave"
                g2, r2 = e.array("indx2"), ref.arr("indx2")
                assert np.array_equal(g2[: rs.nenter], r2[: rs.nenter]), "indx2 entering"
                assert np.array_equal(g2[rs.ileave - 1:], r2[rs.ileave - 1:]), "indx2 leaving"
        if self.status == S.ST_NEED_EVAL:
            z_e, d_e = e.array("z"), e.array("d")
            z_r = ref.arr("z")
            close(z_e, z_r, rtol, "z (subspace point)")
            # d = z - x inherits the error of z, so it is measured on the
            # scale of z (it is often far smaller than z late in a solve)
            zscale = max(float(np.max(np.abs(z_r))), 1e-300)
            derr = float(np.max(np.abs(d_e - ref.arr("d")))) / zscale
            assert derr <= 2 * rtol, f"d = z - x: error {derr:.3e} of max|z|"
            # the line search and the memory update then start from the
            # GPU's z and d on both sides
            z_r[:] = z_e
            ref.arr("d")[:] = d_e
            if rs.nfree != 0 and rs.col != 0 and rs.wrk:
                c2, m = 2 * rs.col, self.m
                wn_g = e.array("wn").reshape(2 * m, 2 * m)[:c2, :c2]
                wn_r = ref.arr("wn").reshape(2 * m, 2 * m)[:c2, :c2]
                close(np.triu(wn_g), np.triu(wn_r), 100 * rtol, "wn")
        # the first lnsrlb call (task = 0): d'd, stpmx, the first step
        ref.fl.lnsrlb_step(ref.s)
        # a non-descent direction (-4) restarts in the tail, as on the GPU
        pending = 1 if (rs.internalinfo != 0 and rs.col > 0) else 0
        assert int(self.es("restarts")) == rs.restarts + pending, \
            ("restarts", self.es("restarts"), rs.restarts, pending)
        if self.status == S.ST_NEED_EVAL:
            scalar_close(self.es("dtd"), rs.dtd, 10 * rtol, "dtd")
            if rs.stpmx < 1e10:
                scalar_close(self.es("stpmx"), rs.stpmx, 10 * rtol, "stpmx")
            else:
                assert self.es("stpmx") >= 1e10, ("stpmx", self.es("stpmx"))

    def search(self):
        """Trials until the oracle's line search stops, with the GPU's f and
        g handed to the oracle at each one."""
        e, ref, rs, rtol = self.eng, self.ref, self.rs, self.rtol
        while rs.internalinfo == 0 and rs.iback < 20 and rs.task == 1:
            assert self.status == S.ST_NEED_EVAL, ("status", self.status)
            if abs(self.es("stp") - rs.stp) > rtol * abs(rs.stp):
                names = ("stp", "f", "fold", "gd", "gdold", "stpmx", "dtd", "iback", "csave",
                         "iter", "col")
                got = {k: (self.es(k), float(getattr(rs, k))) for k in names}
                raise AssertionError(f"line search differs (engine, oracle): {got}; "
                                     f"isave {list(rs.isave)} dsave {list(rs.dsave)}")
            for k in ("iback", "ifun", "nfgv"):
                assert int(self.es(k)) == int(getattr(rs, k)), (k, self.es(k), getattr(rs, k))
            x_t = e.array("xb")
            close(x_t, ref.arr("x"), rtol, "trial point")
            f_t = e.evaluate(1)
            g_t = e.array("gb")
            if self.feed is not None:
                # forced function values (failure-path tests): the same f
                # and g go to both sides
                f_t, g_new = self.feed(self, f_t, x_t, g_t)
                if g_new is not None:
                    g_t = np.ascontiguousarray(g_new, dtype=e.np_dtype)
                    e.set_array("gb", g_t)
            ref.arr("x")[:] = x_t
            ref.arr("g")[:] = g_t
            rs.f = f_t
            ref.fl.lnsrlb_step(ref.s)
            self.counts["trials"] += 1
            self.status = self.run_after_eval(f_t)

    def tail(self):
        """Failure handling, stopping tests and the memory update.  Returns
        the oracle's TAIL_* result."""
        e, ref, rs, rtol, m = self.eng, self.ref, self.rs, self.rtol, self.m
        r = ref.fl.iter_tail(ref.s)
        if r == self.oracle.TAIL_END:
            assert self.status == S.ST_DONE, ("status", self.status, "oracle info", rs.info)
            assert int(self.es("info")) == rs.info, ("info", self.es("info"), rs.info)
            assert int(self.es("iter")) == rs.iter, ("iter", self.es("iter"), rs.iter)
            close(e.array("x"), ref.arr("x"), rtol, "x at the end")
            return r
        if r == self.oracle.TAIL_RESTART:
            assert self.status == S.ST_RESTART, ("status", self.status)
            assert int(self.es("restarts")) == rs.restarts
            assert int(self.es("col")) == 0
            self.counts["restarts"] += 1
            self.status = e.run(S.PROG_ITER_START)
            return r
        # accepted, stopping tests passed, memory updated or skipped
        assert self.status != S.ST_DONE, ("engine stopped", self.es("info"))
        assert int(self.es("iter")) == rs.iter, ("iter", self.es("iter"), rs.iter)
        x_e, g_e = e.array("x"), e.array("g")
        assert self.es("sbgnrm") == self.bulk_sbgnrm(x_e, g_e), "projgr (bit-exact)"
        scalar_close(self.es("sbgnrm"), rs.sbgnrm, rtol, "sbgnrm vs oracle")
        scalar_close(self.es("tf"), rs.tf, 10 * rtol, "|x - x_old|")
        for k in ("updatd", "nskip", "iupdat", "col"):
            assert int(self.es(k)) == int(getattr(rs, k)), (k, self.es(k), getattr(rs, k))
        if int(self.es("restarts")) != rs.restarts:
            # the GPU chain has gone on into the next head and reset there;
            # its memory is no longer this iteration's
            return r
        assert int(self.es("head")) == rs.head - 1, ("head", self.es("head"), rs.head)
        if rs.col > 0:
            assert int(self.es("itail")) == rs.itail - 1, ("itail", self.es("itail"), rs.itail)
        if not rs.updatd:
            self.counts["skips"] += 1
            return r
        col = rs.col
        scalar_close(self.es("theta"), rs.theta, 10 * rtol, "theta")
        scalar_close(self.es("rr"), rs.rr, 10 * rtol, "y'y")
        scalar_close(self.es("dr"), rs.dr, 10 * rtol, "s'y")
        sy_g = e.array("sy").reshape(m, m)[:col, :col]
        sy_r = ref.arr("sy").reshape(m, m)[:col, :col]
        close(np.tril(sy_g), np.tril(sy_r), 10 * rtol, "S'Y")
        ss_g = e.array("ss").reshape(m, m)[:col, :col]
        ss_r = ref.arr("ss").reshape(m, m)[:col, :col]
        close(np.triu(ss_g), np.triu(ss_r), 10 * rtol, "S'S")
        wt_g = e.array("wt").reshape(m, m)[:col, :col]
        wt_r = ref.arr("wt").reshape(m, m)[:col, :col]
        close(np.triu(wt_g), np.triu(wt_r), 100 * rtol, "wt (formt)")
        slots = [(rs.head - 1 + j) % m for j in range(col)]
        n = rs.n
        for name in ("ws", "wy"):
            wg = e.array(name).reshape(n, m)[:, slots]
            wr = ref.arr(name).reshape(n, m)[:, slots]
            close(wg, wr, 10 * rtol, f"{name} ring columns")
        return r

    def iteration(self):
        """One outer iteration of both.  Returns False once the solve ended."""
        while True:
            self.head()
            self.search()
            r = self.tail()
            if r != self.oracle.TAIL_RESTART:
                break
        if r == self.oracle.TAIL_END:
            return False
        self.iters += 1
        return True

    def run(self, max_iters):
        more = self.start()
        while more and self.iters < max_iters:
            more = self.iteration()
        return more

    def close(self):
        self.eng.close()
