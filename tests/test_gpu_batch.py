"""The batched form (lbfgsb_batch_*): many problems in one persistent kernel.

Each problem is solved by one workgroup, or by a gang of workgroups over
segments of its elements, with the iteration chain of the plain form, so
the checks are those of the plain solves -- the minimum against the CPU
oracle's, bounds, info codes -- plus properties of the batched form itself:
a problem's result does not depend on what else is in the batch
(bit-identical alone and in a mix), suspending and resuming with an
iteration budget gives bit-identical results, slots are reused, and bad
input is rejected as lbfgsbminimize rejects it.
"""
import numpy as np
import pytest

from conftest import gpu_available, rel_error, solve_tolerance

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not gpu_available(), reason="needs a ROCm GPU")]


def _mix(seed=7):
    from lbfgsb_amd import problems as pb
    return [("penta_1500", pb.penta_qp(1500, seed=seed), 1e-6, 1000),
            ("penta_4100", pb.penta_qp(4100, seed=seed + 1), 1e-6, 1000),
            ("grid_quad_30x34", pb.grid_quadratic(30, 34, seed=seed + 2), 1e-8, 1000),
            ("grid_cos_40", pb.grid_cos(40, seed=seed + 3), 1e-7, 2000),
            # Rosenbrock runs to its minimiser (f* = 0): stopped early by a
            # loose eps_f, two rounding-different paths stop at different f
            ("rosen_3000", pb.rosenbrock(3000, seed=seed + 4), 1e-13, 3000),
            ("penta_777", pb.penta_qp(777, seed=seed + 5), 1e-6, 1000),
            # first Cauchy rounds of ~10^4 candidates: the bucketed rounds,
            # several groups, sorted bucket by bucket or by the network
            ("penta_24000", pb.penta_qp(24000, seed=seed + 6), 1e-6, 1000)]


def _solve(cases, dtype, threads=256, budget=0, segment=0, max_gang=64, gangs=None):
    from lbfgsb_amd.batch import Batch
    cap = sum(p.n + 64 for _, p, _, _ in cases)
    b = Batch(dtype, cap, len(cases) + 2)
    b.configure(segment, max_gang if segment else 1)
    slots = []
    for name, p, eps, mi in cases:
        s = b.add(p, eps=eps, max_iter=mi)
        assert s >= 0, (name, s)
        slots.append(s)
    if gangs is not None:
        gangs.extend(b.gang(s) for s in slots)
    if budget:
        for _ in range(10000):
            if b.run(iter_budget=budget, threads=threads) == 0:
                break
    else:
        assert b.run(threads=threads) == 0
    f, it, info, done = b.results()
    xs = [b.x(s).double().cpu().numpy() for s in slots]
    out = [(f[s], it[s], info[s], done[s], x) for s, x in zip(slots, xs)]
    b.close()
    return out


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_batch_matches_oracle(dtype, oracle):
    from lbfgsb_amd import problems as pb
    cases = _mix()
    print(f"case batch mix {np.dtype(dtype).name}", flush=True)
    res = _solve(cases, dtype)
    for (name, p, eps, mi), (f, it, info, done, x) in zip(cases, res):
        ref = oracle.Solver(p, dtype=np.float64, eps=eps, max_iter=mi).run()
        f_x, _ = pb.objective_numpy(p, x)
        assert done == 1 and info in (1, 2, 4, 5), (name, info, done)
        lo = np.where((p.nbd == 1) | (p.nbd == 2), p.l, -np.inf)
        hi = np.where((p.nbd == 2) | (p.nbd == 3), p.u, np.inf)
        slack = 1e-12 if dtype == np.float64 else 1e-6
        assert np.all(x >= lo - slack) and np.all(x <= hi + slack), name
        f_ref = ref.st.f
        rel = rel_error(f_x, f_ref, p.kind)
        # the reported f is the objective at the returned x (rounded to the
        # bulk type in fp32, as a callback hands it over)
        rel_rep = rel_error(f, f_x, p.kind)
        tol = solve_tolerance(eps, f_ref, p.kind, 1e-6 if dtype == np.float64 else 1e-5)
        print(name, np.dtype(dtype).name, "batch f", f_x, "oracle f", f_ref, "rel", rel,
              "iters", it, ref.st.iter, "info", info, ref.st.info)
        assert rel < tol, name
        assert rel_rep < (1e-12 if dtype == np.float64 else 1e-6), (name, f, f_x)


def test_batch_alone_equals_in_mix():
    """One problem alone and the same problem among others: bit-identical x,
    f and iteration count (each problem is solved by one workgroup with fixed
    reduction orders)."""
    cases = _mix(seed=11)
    print("case alone-vs-mix", flush=True)
    mix = _solve(cases, np.float32)
    for k in (0, 4):
        alone = _solve([cases[k]], np.float32)[0]
        assert alone[0] == mix[k][0] and alone[1] == mix[k][1] and alone[2] == mix[k][2]
        assert np.array_equal(alone[4], mix[k][4]), cases[k][0]


def test_batch_budget_resume_is_exact():
    """Suspending every problem after 3 iterations and resuming it in the next
    launch gives exactly the result of one uninterrupted launch."""
    cases = _mix(seed=13)[:4]
    print("case budget resume", flush=True)
    full = _solve(cases, np.float64)
    steps = _solve(cases, np.float64, budget=3)
    for a, b in zip(full, steps):
        assert a[0] == b[0] and a[1] == b[1] and a[2] == b[2]
        assert np.array_equal(a[4], b[4])


def test_batch_threads_agree():
    """256- and 512-thread workgroups reach the same minima."""
    cases = _mix(seed=17)[:3]
    print("case threads 256 vs 512", flush=True)
    a = _solve(cases, np.float64, threads=256)
    b = _solve(cases, np.float64, threads=512)
    for (name, p, _, _), ra, rb in zip(cases, a, b):
        assert abs(ra[0] - rb[0]) <= 1e-8 * max(1.0, abs(ra[0])), name


def test_batch_slots_and_bad_input():
    import torch
    from lbfgsb_amd import problems as pb
    from lbfgsb_amd.batch import Batch
    print("case slots and bad input", flush=True)
    p1 = pb.penta_qp(1000, seed=3)
    p2 = pb.penta_qp(900, seed=4)
    b = Batch(np.float64, 2048, 2)
    s1 = b.add(p1, eps=1e-6)
    assert s1 >= 0
    # m > n and an unknown objective kind are rejected
    bad = pb.penta_qp(5, seed=5)
    assert b.add(bad, eps=1e-6) == -1
    t = torch.zeros(1000, dtype=torch.float64, device="cuda")
    nb = torch.zeros(1000, dtype=torch.int32, device="cuda")
    assert b.add_device(1000, t, nb, t, t, kind=9) == -1
    # no room for a second 1000-variable problem beside the first in 2048
    # elements? 1024 + 1024 fit; a third does not
    s2 = b.add(p2, eps=1e-6)
    assert s2 >= 0
    assert b.add(pb.penta_qp(100, seed=6), eps=1e-6) == -3
    assert b.run() == 0
    f, it, info, done = b.results()
    assert done[s1] == 1 and done[s2] == 1
    b.remove(s1)
    s3 = b.add(pb.penta_qp(1000, seed=3), eps=1e-6)
    assert s3 == s1
    assert b.run() == 0
    f2, it2, info2, done2 = b.results()
    # the same problem in a reused slot and range: the same result
    assert f2[s3] == f[s1] and it2[s3] == it[s1]
    b.close()


def _gang_mix(seed=21):
    from lbfgsb_amd import problems as pb
    return [("penta_20000", pb.penta_qp(20000, seed=seed), 1e-6, 1000),
            ("grid_quad_100x120", pb.grid_quadratic(100, 120, seed=seed + 1), 1e-8, 1000),
            ("grid_cos_90", pb.grid_cos(90, seed=seed + 2), 1e-7, 2000),
            ("rosen_10000", pb.rosenbrock(10000, seed=seed + 3), 1e-13, 3000),
            ("penta_4100", pb.penta_qp(4100, seed=seed + 4), 1e-6, 1000),
            ("penta_1500", pb.penta_qp(1500, seed=seed + 5), 1e-6, 1000)]


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_batch_gangs_match_oracle(dtype, oracle):
    """Gangs of workgroups (4096-element segments: 2 to 5 members here)
    reach the oracle's minima, like one workgroup per problem."""
    from lbfgsb_amd import problems as pb
    cases = _gang_mix()
    print(f"case gang mix {np.dtype(dtype).name}", flush=True)
    gangs = []
    res = _solve(cases, dtype, segment=4096, gangs=gangs)
    assert gangs == [5, 3, 2, 3, 2, 1], gangs
    for (name, p, eps, mi), (f, it, info, done, x) in zip(cases, res):
        ref = oracle.Solver(p, dtype=np.float64, eps=eps, max_iter=mi).run()
        f_x, _ = pb.objective_numpy(p, x)
        assert done == 1 and info in (1, 2, 4, 5), (name, info, done)
        lo = np.where((p.nbd == 1) | (p.nbd == 2), p.l, -np.inf)
        hi = np.where((p.nbd == 2) | (p.nbd == 3), p.u, np.inf)
        slack = 1e-12 if dtype == np.float64 else 1e-6
        assert np.all(x >= lo - slack) and np.all(x <= hi + slack), name
        f_ref = ref.st.f
        rel = rel_error(f_x, f_ref, p.kind)
        rel_rep = rel_error(f, f_x, p.kind)
        tol = solve_tolerance(eps, f_ref, p.kind, 1e-6 if dtype == np.float64 else 1e-5)
        print(name, np.dtype(dtype).name, "gang f", f_x, "oracle f", f_ref, "rel", rel,
              "iters", it, ref.st.iter, "info", info, ref.st.info)
        assert rel < tol, name
        assert rel_rep < (1e-12 if dtype == np.float64 else 1e-6), (name, f, f_x)


def test_batch_gangs_alone_equals_in_mix():
    """A gang-solved problem alone and in a mix: bit-identical (the gang size
    depends on n and the configuration, the reductions combine in member
    order)."""
    cases = _gang_mix(seed=31)
    print("case gang alone-vs-mix", flush=True)
    mix = _solve(cases, np.float32, segment=4096)
    for k in (0, 3):
        alone = _solve([cases[k]], np.float32, segment=4096)[0]
        assert alone[0] == mix[k][0] and alone[1] == mix[k][1] and alone[2] == mix[k][2]
        assert np.array_equal(alone[4], mix[k][4]), cases[k][0]


def test_batch_gangs_budget_resume_is_exact():
    """Gangs suspended every 4 iterations and resumed: the uninterrupted
    result exactly."""
    cases = _gang_mix(seed=41)[:3]
    print("case gang budget resume", flush=True)
    full = _solve(cases, np.float64, segment=4096)
    steps = _solve(cases, np.float64, segment=4096, budget=4)
    for a, b in zip(full, steps):
        assert a[0] == b[0] and a[1] == b[1] and a[2] == b[2]
        assert np.array_equal(a[4], b[4])


def test_batch_gang_wider_than_launch_is_rejected():
    from lbfgsb_amd import problems as pb
    from lbfgsb_amd.batch import Batch
    print("case gang wider than launch", flush=True)
    b = Batch(np.float64, 20000 + 64, 2)
    b.configure(4096, 64)
    assert b.add(pb.penta_qp(20000, seed=2), eps=1e-6) >= 0
    with pytest.raises(ValueError):
        b.run(blocks=2)
    with pytest.raises(ValueError):
        b.configure(1000, 4)
    b.close()
