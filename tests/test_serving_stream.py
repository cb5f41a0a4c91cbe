"""The serving workload's request stream (host side, no GPU): deterministic,
with the stated distributions and shapes every generator accepts."""
import math

import numpy as np

from lbfgsb_amd import problems as pb


def test_stream_is_deterministic_and_shaped():
    a = pb.serving_stream(2000, qps=8.0)
    b = pb.serving_stream(2000, qps=8.0)
    assert a == b
    n = np.array([r["n"] for r in a])
    assert n.min() >= 900 and n.max() <= 4_000_000
    # log-uniform: about half the sizes below the geometric mean of the range
    gm = math.sqrt(1e3 * 4e6)
    assert 0.45 < np.mean(n < gm) < 0.55
    kinds = np.bincount([r["kind"] for r in a], minlength=5)[1:]
    assert kinds.min() > 400 and kinds.sum() == 2000
    fp32 = np.mean([r["dtype"] == 0 for r in a])
    assert 0.66 < fp32 < 0.74
    arr = np.array([r["arrival"] for r in a])
    assert np.all(np.diff(arr) > 0)
    # Poisson at 8 per second: the mean gap is 1/8 s
    assert abs(np.mean(np.diff(arr)) - 0.125) < 0.01
    for r in a:
        if r["kind"] in (pb.OBJ_GRID_QUAD, pb.OBJ_GRID_COS):
            assert r["nx"] * r["ny"] == r["n"]
        if r["kind"] == pb.OBJ_ROSEN:
            assert r["n"] % 2 == 0
        assert r["eps"] == pb.SERVE_EPS
        assert r["max_iteration"] == pb.KIND_SETTINGS[r["kind"]][1]
    assert len({r["seed"] for r in a}) == 2000


def test_problem_of_builds_the_request():
    for r in pb.serving_stream(40, qps=8.0, hi=5e4):
        p = pb.problem_of(r)
        assert p.n == r["n"] and p.kind == r["kind"]
