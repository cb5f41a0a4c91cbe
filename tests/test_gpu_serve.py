"""The device problem generator, graph-replayed batch launches and the
serving loop (lbfgsb_generate, lbfgsb_batch_use_graph, lbfgsb_serve_stream).

The generator is checked against the host generator of the same problems
(uniform-derived arrays bit-identical, normal draws to a few ulps).  Graph
replays must give the eager launches' results bit for bit.  The serving
loop must give every request the result it has alone (bit-identical f,
iteration count and info), whatever else arrives with it, keep what it
holds within the budget, and reach the oracle's minima.
"""
import ctypes as C

import numpy as np
import pytest

from conftest import gpu_available, rel_error, solve_tolerance

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not gpu_available(), reason="needs a ROCm GPU")]


def _generate(kind, n, nx, ny, seed, dtype):
    import torch
    from lbfgsb_amd import _lib
    tdt = torch.float64 if dtype == np.float64 else torch.float32
    arrs = {k: torch.zeros(n, dtype=tdt, device="cuda")
            for k in ("x0", "l", "u", "b", "c", "a0", "a1", "a2")}
    nbd = torch.zeros(n, dtype=torch.int32, device="cuda")
    ptr = lambda t: C.c_void_p(t.data_ptr())
    stream = C.c_void_p(torch.cuda.current_stream().cuda_stream)
    rc = _lib.lib().lbfgsb_generate(1 if dtype == np.float64 else 0, kind, n, nx, ny, seed,
                                    ptr(arrs["x0"]), ptr(nbd), ptr(arrs["l"]), ptr(arrs["u"]),
                                    ptr(arrs["b"]), ptr(arrs["c"]), ptr(arrs["a0"]),
                                    ptr(arrs["a1"]), ptr(arrs["a2"]), stream)
    assert rc == 0
    out = {k: v.cpu().numpy() for k, v in arrs.items()}
    out["nbd"] = nbd.cpu().numpy()
    return out


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_generator_matches_host(dtype):
    from lbfgsb_amd import problems as pb
    cases = [pb.grid_quadratic(37, 41, seed=101), pb.grid_cos(43, 29, seed=102),
             pb.rosenbrock(2000, seed=103), pb.penta_qp(3001, seed=104)]
    seeds = [101, 102, 103, 104]
    for p, seed in zip(cases, seeds):
        print(f"case generate {p.name} {np.dtype(dtype).name}", flush=True)
        g = _generate(p.kind, p.n, p.nx, p.ny, seed, dtype)
        q = p.cast(dtype)
        assert np.array_equal(g["nbd"], q.nbd), p.name
        for k in ("x0", "l", "u", "c", "a0", "a1", "a2"):
            ref = getattr(q, k)
            if ref is not None:
                assert np.array_equal(g[k], ref), (p.name, k)
        if q.b is not None:
            if p.kind == pb.OBJ_GRID_QUAD:
                assert np.array_equal(g["b"], q.b), p.name
            else:
                # Box-Muller through the device's log and cos: a few ulps
                ulp = np.spacing(np.abs(q.b).astype(dtype)).astype(np.float64)
                err = np.abs(g["b"].astype(np.float64) - q.b.astype(np.float64))
                assert np.all(err <= 4 * ulp), (p.name, float((err / ulp).max()))


def test_generator_rejects_bad_specs():
    import torch
    from lbfgsb_amd import _lib
    L = _lib.lib()
    t = torch.zeros(64, dtype=torch.float64, device="cuda")
    nb = torch.zeros(64, dtype=torch.int32, device="cuda")
    p = C.c_void_p(t.data_ptr())
    q = C.c_void_p(nb.data_ptr())
    s = C.c_void_p(torch.cuda.current_stream().cuda_stream)
    assert L.lbfgsb_generate(1, 1, 64, 7, 9, 1, p, q, p, p, p, p, p, p, p, s) == -1  # 7*9 != 64
    assert L.lbfgsb_generate(1, 3, 63, 0, 0, 1, p, q, p, p, p, p, p, p, p, s) == -1  # odd n
    assert L.lbfgsb_generate(1, 9, 64, 0, 0, 1, p, q, p, p, p, p, p, p, p, s) == -1  # kind
    assert L.lbfgsb_generate(1, 4, 64, 0, 0, 1, p, q, p, p, p, None, None, None, None,
                             s) == -1  # penta without its arrays


def _mixed_specs():
    from lbfgsb_amd import problems as pb
    return [(pb.OBJ_PENTA, 20000, 0, 0, 7001), (pb.OBJ_GRID_QUAD, 3000, 50, 60, 7002),
            (pb.OBJ_GRID_COS, 1600, 40, 40, 7003), (pb.OBJ_ROSEN, 9000, 0, 0, 7004),
            (pb.OBJ_PENTA, 1500, 0, 0, 7005)]


def _batch_run(dtype, specs, graph, budget, segment=4096):
    from lbfgsb_amd import problems as pb
    from lbfgsb_amd.batch import Batch
    b = Batch(dtype, sum(n + 64 for _, n, _, _, _ in specs), len(specs))
    b.configure(segment, 64 if segment else 1)
    b.use_graph(graph)
    slots = []
    for kind, n, nx, ny, seed in specs:
        eps, mi = pb.KIND_SETTINGS[kind]
        s = b.add_generated(kind, n, nx, ny, seed, eps=eps, max_iter=mi)
        assert s >= 0
        slots.append(s)
    launches = 0
    while True:
        left = b.run(iter_budget=budget)
        launches += 1
        if left == 0 or launches > 5000:
            break
    f, it, info, done = b.results()
    xs = [b.x(s).double().cpu().numpy() for s in slots]
    b.close()
    return [(f[s], it[s], info[s], done[s]) for s in slots], xs, launches


def test_graph_replays_match_eager_launches():
    """Budgeted launches replayed from one captured graph (many replays with
    the host's uploads and result reads between them) against the same
    launches made eagerly: bit-identical."""
    specs = _mixed_specs()
    print("case graph vs eager", flush=True)
    eager, xe, n1 = _batch_run(np.float32, specs, False, 7)
    graph, xg, n2 = _batch_run(np.float32, specs, True, 7)
    assert n1 == n2 and n1 > 5
    for a, b in zip(eager, graph):
        assert a == b
    for a, b in zip(xe, xg):
        assert np.array_equal(a, b)


def test_generated_problem_solves_like_uploaded(oracle):
    """A generated problem reaches the oracle's minimum of the host-built
    problem (the normal draws may differ in the last bits)."""
    from lbfgsb_amd import problems as pb
    specs = _mixed_specs()
    print("case generated vs oracle", flush=True)
    res, xs, _ = _batch_run(np.float64, specs, True, 0)
    for (kind, n, nx, ny, seed), (f, it, info, done), x in zip(specs, res, xs):
        req = dict(kind=kind, n=n, nx=nx, ny=ny, seed=seed)
        p = pb.problem_of(req)
        eps, mi = pb.KIND_SETTINGS[kind]
        ref = oracle.Solver(p, dtype=np.float64, eps=eps, max_iter=mi).run()
        f_x, _ = pb.objective_numpy(p, x)
        rel = rel_error(f_x, ref.st.f, p.kind)
        print(p.name, "f", f_x, "oracle", ref.st.f, "rel", rel, "iters", it, ref.st.iter)
        assert done == 1 and info in (1, 2, 4, 5)
        assert rel < solve_tolerance(eps, ref.st.f, kind, 1e-6), p.name


def _stream(count=14, seed=31):
    from lbfgsb_amd import problems as pb
    reqs = pb.serving_stream(count, qps=400.0, lo=1e3, hi=3e4, seed=seed)
    return reqs


def _serve(reqs, **kw):
    from lbfgsb_amd.serve import serve
    args = dict(budget_bytes=2e9, max_slots=64, segment=8192, quantum=6)
    args.update(kw)
    return serve(reqs, **args)


def test_serve_alone_equals_in_mix():
    """Each request's answer in a continuously batched stream is the answer
    it gets served alone: bit-identical f, iterations and info."""
    reqs = _stream()
    print("case serve mix", flush=True)
    mix, st = _serve(reqs)
    assert st["max_resident"] >= 2
    assert all(r["info"] in (1, 2, 4, 5) for r in mix), [r["info"] for r in mix]
    assert any(r["gang"] > 1 for r in mix)
    for k in range(0, len(reqs), 3):
        one = dict(reqs[k], arrival=0.0)
        print(f"case serve alone {k}", flush=True)
        alone, _ = _serve([one])
        a, b = alone[0], mix[k]
        assert (a["f"], a["iter"], a["info"], a["gang"]) == (b["f"], b["iter"], b["info"],
                                                              b["gang"]), k
        assert b["t_admit"] >= b["t_arrival"] and b["t_done"] >= b["t_admit"]


def test_serve_admission_respects_budget():
    """A budget that holds only part of the stream at once: requests wait
    for room, the server holds no more than the budget, and the answers are
    those of an unconstrained run."""
    from lbfgsb_amd import _lib
    reqs = _stream(count=10, seed=37)
    print("case serve budget", flush=True)
    big, _ = _serve(reqs)
    L = _lib.lib()
    # room for about 40k fp32 and 20k fp64 elements at a time
    budget = L.lbfgsb_batch_bytes(0, 40000, 4) + L.lbfgsb_batch_bytes(1, 20000, 4)
    share = L.lbfgsb_batch_bytes(0, 40000, 4) / budget
    small, st = _serve(reqs, budget_bytes=budget, fp32_share=share, max_slots=4)
    assert st["arena_bytes"] <= budget
    assert st["max_waiting"] >= 1
    for a, b in zip(big, small):
        if b["info"] == -3:
            continue  # larger than its arena under this budget
        assert (a["f"], a["iter"], a["info"]) == (b["f"], b["iter"], b["info"])
    waited = [b["t_admit"] - b["t_arrival"] for b in small if b["info"] >= 0]
    assert max(waited) > 0


@pytest.mark.parametrize("polish", [False, True], ids=["plain", "polish"])
def test_serve_reaches_oracle_minima(oracle, polish):
    """Every served answer is as good as the FP64 oracle's solve of the same
    request, up to the resolution of the decrease test both stop on
    (solve_tolerance): (f - f_oracle) / |f_oracle| below it, one-sided,
    since two solves of an ill-conditioned grid quadratic that stop on an
    absolute decrease of 1e-8 can both be 1e-4 above its minimum and either
    one the lower.  fp32 answers without the polish have the float iterate's
    resolution as well (1e-5); polished ones are FP64 answers."""
    from lbfgsb_amd import problems as pb
    reqs = _stream(count=8, seed=41)
    print(f"case serve vs oracle polish={polish}", flush=True)
    res, _ = _serve(reqs, polish=polish)
    for q, r in zip(reqs, res):
        p = pb.problem_of(q)
        ref = oracle.Solver(p, dtype=np.float64, eps=q["eps"], max_iter=q["max_iteration"]).run()
        fp64_answer = q["dtype"] == 1 or polish
        assert r["polished"] == int(polish and q["dtype"] == 0)
        if r["polished"]:
            assert r["iter_fp32"] > 0 and r["iter"] >= r["iter_fp32"]
        floor = 1e-6 if fp64_answer else 1e-5
        tol = solve_tolerance(q["eps"], ref.st.f, p.kind, floor)
        worse = (r["f"] - ref.st.f) / (1.0 if p.kind == pb.OBJ_ROSEN else abs(ref.st.f))
        print(p.name, q["dtype"], "served f", r["f"], "oracle", ref.st.f, "above by", worse,
              "bound", tol, "iters", r["iter"], r["iter_fp32"], ref.st.iter, flush=True)
        assert r["info"] in (1, 2, 4, 5)
        assert worse < tol, p.name


def test_serve_polish_alone_equals_in_mix():
    """With the FP64 polish an fp32 request's answer is still the one it
    gets served alone (both phases are functions of the request)."""
    reqs = [dict(r, dtype=0) for r in _stream(count=6, seed=43)]
    print("case serve polish mix", flush=True)
    mix, _ = _serve(reqs, polish=True)
    for k in (0, 3, 5):
        alone, _ = _serve([dict(reqs[k], arrival=0.0)], polish=True)
        a, b = alone[0], mix[k]
        assert a["polished"] == b["polished"] == 1
        assert (a["f"], a["iter"], a["info"], a["iter_fp32"]) == \
            (b["f"], b["iter"], b["info"], b["iter_fp32"]), k


def test_serve_rejects_what_cannot_run():
    from lbfgsb_amd import problems as pb
    reqs = [dict(kind=pb.OBJ_GRID_QUAD, dtype=0, n=100, nx=7, ny=9, seed=1, eps=1e-6,
                 max_iteration=10, arrival=0.0),
            dict(kind=pb.OBJ_PENTA, dtype=1, n=10**6, nx=0, ny=0, seed=2, eps=1e-6,
                 max_iteration=10, arrival=0.0),
            dict(kind=pb.OBJ_PENTA, dtype=0, n=2000, nx=0, ny=0, seed=3, eps=1e-6,
                 max_iteration=1000, arrival=0.001)]
    print("case serve rejects", flush=True)
    res, _ = _serve(reqs, budget_bytes=2e8)
    assert res[0]["info"] == -1       # nx * ny != n
    assert res[1]["info"] == -3       # larger than its arena
    assert res[2]["info"] in (1, 2, 4, 5)
