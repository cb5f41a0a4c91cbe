"""Deterministic synthetic problem generators for the benchmark workloads.

Every random draw comes from one counter-based generator (splitmix64 of
``seed ^ stream*phi64 ^ i``), so a value depends only on (seed, stream, i):
the same arrays come out on any host, at any size, in any order of
generation.  Streams: 1 = b, 2 = c, 3 = x0, 4 = bound offsets, 5 = nbd draw,
6 = A, 7 = n draw, 8 = arrivals, 9 = request mix.

Workloads (n, bounds and objectives):

* ``grid_quadratic`` -- n = nx*ny, h = 1/(nx+1).
  f(v) = 0.5*sum_edges (v_a - v_b)^2 - h^2 * sum b_i v_i over the 5-point
  grid with a zero Dirichlet ring (that is 0.5*h^2*|grad_h v|^2 with forward
  differences), b ~ U(0,1), nbd = 2, l = -0.05, u = 0.05, x0 = 0.
* ``grid_cos`` -- f(v) = 0.5*sum_edges (dv)^2 + h^2*sum(c_i cos v_i - b_i v_i),
  c ~ U(0.5, 2), b ~ N(0,1), nbd in {0,1,2} w.p. 0.2/0.3/0.5,
  x0 ~ U(-0.5, 0.5), l = x0 - U(0.1,1), u = x0 + U(0.1,1).
* ``rosenbrock`` -- extended Rosenbrock over pairs, nbd = 2, bounds
  [-1.5, 2] on 1-based odd indices and [-0.5, 1.5] on even ones,
  x0 ~ U(-1.2, 1.2); the minimiser x* = 1 is interior.
* ``penta_qp`` -- f = 0.5 x'Ax - b'x, A symmetric pentadiagonal with
  diagonal U(4,8) and both off-diagonals U(-1,1)*0.9 (strictly diagonally
  dominant, hence SPD), b ~ N(0,1), l ~ U(-1,-0.1), u ~ U(0.1,1), x0 = 0.
* ``dsscfg`` -- the reference's example problem (MINPACK-2 steady-state
  combustion, examples/dsscfg/dsscfg_cpu.cpp): piecewise-linear elements on
  two triangles per grid point, hx = 1/(nx+1), hy = 1/(ny+1),
  f = area*(0.5*sum|grad v|^2 - lambda*sum trapezoid(exp v)), unconstrained
  (nbd = 0), the reference's standard start ('XS').
"""
from __future__ import annotations

import dataclasses
import math

import numpy as np

PHI64 = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)

SEED_BASE = 0x4C42_4647_0000

STREAM_B, STREAM_C, STREAM_X0, STREAM_BOUND, STREAM_NBD = 1, 2, 3, 4, 5
STREAM_A, STREAM_N, STREAM_ARRIVAL, STREAM_MIX = 6, 7, 8, 9

# objective kinds shared with the C-ABI (include/lbfgsb_amd.h)
OBJ_GRID_QUAD, OBJ_GRID_COS, OBJ_ROSEN, OBJ_PENTA = 1, 2, 3, 4
OBJ_DSSCFG = 6


def splitmix64(z: np.ndarray) -> np.ndarray:
    z = z + PHI64
    z = (z ^ (z >> np.uint64(30))) * _M1
    z = (z ^ (z >> np.uint64(27))) * _M2
    return z ^ (z >> np.uint64(31))


def uniform(seed: int, stream: int, n: int, offset: int = 0) -> np.ndarray:
    """u in [0, 1) with 53 random bits, one per counter value."""
    with np.errstate(over="ignore"):
        key = np.uint64(seed) ^ (np.uint64(stream) * PHI64)
        i = np.arange(offset, offset + n, dtype=np.uint64)
        z = splitmix64(key ^ i)
    return (z >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


def normal(seed: int, stream: int, n: int) -> np.ndarray:
    """Box-Muller on counter pairs (2i, 2i+1); u1 is taken as 1-u so it is
    never 0."""
    u = uniform(seed, stream, 2 * n)
    u1 = 1.0 - u[0::2]
    u2 = u[1::2]
    return np.sqrt(-2.0 * np.log(u1)) * np.cos(2.0 * math.pi * u2)


@dataclasses.dataclass
class Problem:
    """A bound-constrained problem instance (float64 master copy)."""
    kind: int
    n: int
    x0: np.ndarray
    nbd: np.ndarray
    l: np.ndarray
    u: np.ndarray
    nx: int = 0
    ny: int = 0
    h: float = 0.0
    b: np.ndarray | None = None
    c: np.ndarray | None = None
    a0: np.ndarray | None = None
    a1: np.ndarray | None = None
    a2: np.ndarray | None = None
    name: str = ""

    def cast(self, dtype) -> "Problem":
        f = lambda a: None if a is None else np.ascontiguousarray(a, dtype=dtype)
        return dataclasses.replace(
            self, x0=f(self.x0), l=f(self.l), u=f(self.u), b=f(self.b),
            c=f(self.c), a0=f(self.a0), a1=f(self.a1), a2=f(self.a2),
            nbd=np.ascontiguousarray(self.nbd, dtype=np.int32))


def grid_quadratic(nx: int, ny: int | None = None, seed: int = SEED_BASE + 0) -> Problem:
    ny = nx if ny is None else ny
    n = nx * ny
    return Problem(
        kind=OBJ_GRID_QUAD, n=n, nx=nx, ny=ny, h=1.0 / (nx + 1),
        x0=np.zeros(n), nbd=np.full(n, 2, dtype=np.int32),
        l=np.full(n, -0.05), u=np.full(n, 0.05),
        b=uniform(seed, STREAM_B, n), name=f"grid_quadratic_{nx}x{ny}")


def grid_cos(nx: int, ny: int | None = None, seed: int = SEED_BASE + 1) -> Problem:
    ny = nx if ny is None else ny
    n = nx * ny
    x0 = uniform(seed, STREAM_X0, n) - 0.5
    off = uniform(seed, STREAM_BOUND, 2 * n)
    draw = uniform(seed, STREAM_NBD, n)
    nbd = np.where(draw < 0.2, 0, np.where(draw < 0.5, 1, 2)).astype(np.int32)
    lo = x0 - (0.1 + 0.9 * off[0::2])
    hi = x0 + (0.1 + 0.9 * off[1::2])
    return Problem(
        kind=OBJ_GRID_COS, n=n, nx=nx, ny=ny, h=1.0 / (nx + 1), x0=x0, nbd=nbd,
        l=lo, u=hi, b=normal(seed, STREAM_B, n),
        c=0.5 + 1.5 * uniform(seed, STREAM_C, n), name=f"grid_cos_{nx}x{ny}")


def rosenbrock(n: int, seed: int = SEED_BASE + 2) -> Problem:
    if n % 2:
        raise ValueError("extended Rosenbrock needs an even n")
    odd = (np.arange(n) % 2) == 0  # 1-based odd indices
    return Problem(
        kind=OBJ_ROSEN, n=n, x0=-1.2 + 2.4 * uniform(seed, STREAM_X0, n),
        nbd=np.full(n, 2, dtype=np.int32),
        l=np.where(odd, -1.5, -0.5), u=np.where(odd, 2.0, 1.5),
        name=f"rosenbrock_{n}")


def penta_qp(n: int, seed: int = SEED_BASE + 3) -> Problem:
    a = uniform(seed, STREAM_A, 3 * n)
    a0 = 4.0 + 4.0 * a[0::3]
    a1 = (2.0 * a[1::3] - 1.0) * 0.9
    a2 = (2.0 * a[2::3] - 1.0) * 0.9
    a1[n - 1:] = 0.0
    a2[max(n - 2, 0):] = 0.0
    off = uniform(seed, STREAM_BOUND, 2 * n)
    return Problem(
        kind=OBJ_PENTA, n=n, x0=np.zeros(n), nbd=np.full(n, 2, dtype=np.int32),
        l=-(0.1 + 0.9 * off[0::2]), u=0.1 + 0.9 * off[1::2],
        b=normal(seed, STREAM_B, n), a0=a0, a1=a1, a2=a2, name=f"penta_qp_{n}")


def dsscfg(nx: int, ny: int | None = None, lam: float = 1.0) -> Problem:
    """The reference example's problem: nx*ny unknowns, no bounds, starting
    point x_k = lambda/(lambda+1) * sqrt(min(min(i, nx-i)*hx,
    min(j, ny-j)*hy)) for k = nx*j + i (examples/dsscfg/dsscfg_cpu.cpp:25-37).
    The lambda of the objective travels in h."""
    ny = nx if ny is None else ny
    n = nx * ny
    hx, hy = 1.0 / (nx + 1), 1.0 / (ny + 1)
    i = np.arange(nx)
    j = np.arange(ny)
    tx = np.minimum(i, nx - i) * hx
    ty = np.minimum(j, ny - j) * hy
    x0 = (lam / (lam + 1.0)) * np.sqrt(np.minimum(tx[None, :], ty[:, None]))
    return Problem(kind=OBJ_DSSCFG, n=n, nx=nx, ny=ny, h=lam, x0=x0.reshape(-1),
                   nbd=np.zeros(n, dtype=np.int32), l=np.zeros(n), u=np.zeros(n),
                   name=f"dsscfg_{nx}x{ny}")


def batch_sizes(count: int, lo: float, hi: float, seed: int) -> np.ndarray:
    """n drawn log-uniform in [lo, hi] (floor), one per problem."""
    u = uniform(seed, STREAM_N, count)
    return np.floor(np.exp(math.log(lo) + u * (math.log(hi) - math.log(lo)))).astype(np.int64)


def objective_numpy(p: Problem, x: np.ndarray):
    """Reference float64 evaluation of f and g (used by tests and by the
    projected Gauss-Seidel check)."""
    x = np.asarray(x, dtype=np.float64)
    if p.kind in (OBJ_GRID_QUAD, OBJ_GRID_COS):
        v = x.reshape(p.ny, p.nx)
        pad = np.zeros((p.ny + 2, p.nx + 2))
        pad[1:-1, 1:-1] = v
        lap = 4 * v - pad[1:-1, :-2] - pad[1:-1, 2:] - pad[:-2, 1:-1] - pad[2:, 1:-1]
        ex = np.diff(pad[1:-1, :], axis=1)
        ey = np.diff(pad[:, 1:-1], axis=0)
        h2 = p.h * p.h
        b = p.b.reshape(p.ny, p.nx)
        f = 0.5 * (np.sum(ex * ex) + np.sum(ey * ey)) - h2 * np.sum(b * v)
        g = lap - h2 * b
        if p.kind == OBJ_GRID_COS:
            c = p.c.reshape(p.ny, p.nx)
            f += h2 * np.sum(c * np.cos(v))
            g = g - h2 * c * np.sin(v)
        return float(f), g.reshape(-1)
    if p.kind == OBJ_ROSEN:
        a, b = x[0::2], x[1::2]
        t1, t2 = b - a * a, 1.0 - a
        g = np.empty_like(x)
        g[0::2] = -400.0 * t1 * a - 2.0 * t2
        g[1::2] = 200.0 * t1
        return float(np.sum(100.0 * t1 * t1 + t2 * t2)), g
    if p.kind == OBJ_PENTA:
        ax = penta_matvec(p, x)
        return float(0.5 * x @ ax - p.b @ x), ax - p.b
    if p.kind == OBJ_DSSCFG:
        return dsscfg_numpy(p, x)
    raise ValueError(p.kind)


def dsscfg_numpy(p: Problem, x: np.ndarray):
    """f and g of the dsscfg energy by automatic summation over the two
    triangles of every grid point (the element form, independent of the
    gather order the kernels use)."""
    nx, ny, lam = p.nx, p.ny, p.h
    hx, hy = 1.0 / (nx + 1), 1.0 / (ny + 1)
    area = 0.5 * hx * hy
    pad = np.zeros((ny + 2, nx + 2))
    pad[1:-1, 1:-1] = np.asarray(x, dtype=np.float64).reshape(ny, nx)
    v = pad[1:-1, 1:-1]
    vr, vt = pad[1:-1, 2:], pad[2:, 1:-1]
    vl, vb = pad[1:-1, :-2], pad[:-2, 1:-1]
    dxl, dyl = (vr - v) / hx, (vt - v) / hy
    dxu, dyu = (v - vl) / hx, (v - vb) / hy
    f = area * (0.5 * np.sum(dxl ** 2 + dyl ** 2 + dxu ** 2 + dyu ** 2)
                - lam * np.sum(np.exp(v) + np.exp(vr) + np.exp(vt)
                               + np.exp(vb) + np.exp(vl) + np.exp(v)) / 3.0)
    # gradient of each triangle's terms, scattered onto the padded grid
    gp = np.zeros_like(pad)
    c = slice(1, -1)
    gp[c, c] += -dxl / hx - dyl / hy + dxu / hx + dyu / hy - 2 * lam * np.exp(v) / 3.0
    gp[c, 2:] += dxl / hx - lam * np.exp(vr) / 3.0
    gp[2:, c] += dyl / hy - lam * np.exp(vt) / 3.0
    gp[c, :-2] += -dxu / hx - lam * np.exp(vl) / 3.0
    gp[:-2, c] += -dyu / hy - lam * np.exp(vb) / 3.0
    return float(f), (area * gp[c, c]).reshape(-1)


def penta_matvec(p: Problem, x: np.ndarray) -> np.ndarray:
    ax = p.a0 * x
    ax[:-1] += p.a1[:-1] * x[1:]
    ax[1:] += p.a1[:-1] * x[:-1]
    ax[:-2] += p.a2[:-2] * x[2:]
    ax[2:] += p.a2[:-2] * x[:-2]
    return ax


def batched_qp_specs(count: int = 4096, lo: float = 1e3, hi: float = 2e5,
                     seed: int = SEED_BASE + 3) -> list:
    """Specs of the batched workload: `count` pentadiagonal box QPs, n
    log-uniform in [lo, hi] (stream 7 of `seed`), problem k seeded with
    seed + (k+1)*2^20, tolerance 1e-6, at most 1000 iterations (the request
    dicts of serving_stream, without arrival times)."""
    sizes = batch_sizes(count, lo, hi, seed)
    return [dict(kind=OBJ_PENTA, dtype=0, n=int(n), nx=0, ny=0, seed=seed + ((k + 1) << 20),
                 eps=1e-6, max_iteration=1000) for k, n in enumerate(sizes)]


def batched_qps(count: int = 4096, lo: float = 1e3, hi: float = 2e5,
                seed: int = SEED_BASE + 3) -> list:
    """The batched workload as host problems (see batched_qp_specs)."""
    return [penta_qp(r["n"], seed=r["seed"]) for r in batched_qp_specs(count, lo, hi, seed)]


# Tolerances and iteration caps of the generators' own workloads, used by
# the serving stream: (eps, max_iteration) per objective kind.
KIND_SETTINGS = {OBJ_GRID_QUAD: (1e-8, 1000), OBJ_GRID_COS: (1e-7, 2000),
                 OBJ_ROSEN: (1e-7, 1000), OBJ_PENTA: (1e-6, 1000)}


def request_shape(kind: int, n: int):
    """(n, nx, ny) of a request of about n variables: grids take
    nx = floor(sqrt(n)), ny = n // nx; Rosenbrock an even n."""
    if kind in (OBJ_GRID_QUAD, OBJ_GRID_COS):
        nx = max(1, int(math.isqrt(int(n))))
        ny = max(1, int(n) // nx)
        return nx * ny, nx, ny
    if kind == OBJ_ROSEN:
        n = int(n) - int(n) % 2
        return max(n, 2), 0, 0
    return int(n), 0, 0


# Tolerance of the serving workload: its answers are held to a relative
# objective error of 1e-6 against FP64, which the generators' own
# tolerances do not resolve -- the decrease test is absolute below |f| = 1
# (eps * max(|f|, 1)), and the grid quadratics have |f| ~ 4e-3 -- so every
# request is solved to 1e-10, with its generator's iteration cap.
SERVE_EPS = 1e-10


def serving_stream(count: int = 20000, qps: float = 10.0, lo: float = 1e3, hi: float = 4e6,
                   fp32_frac: float = 0.7, seed: int = SEED_BASE + 4,
                   eps: float | None = SERVE_EPS) -> list:
    """The serving workload: `count` requests with Poisson arrivals at `qps`
    per second (stream 8), n log-uniform in [lo, hi] (stream 7), the
    objective kind uniform over the four generators and the precision fp32
    with probability fp32_frac (stream 9, counters 2k and 2k+1), and problem
    k seeded with seed + (k+1)*2^20.  Each request has tolerance `eps`
    (None: its generator's own) and its generator's iteration cap
    (KIND_SETTINGS)."""
    u = uniform(seed, STREAM_ARRIVAL, count)
    arrivals = np.cumsum(-np.log1p(-u) / qps)
    sizes = batch_sizes(count, lo, hi, seed)
    mix = uniform(seed, STREAM_MIX, 2 * count)
    kinds = 1 + np.minimum((mix[0::2] * 4).astype(np.int64), 3)
    out = []
    for k in range(count):
        kind = int(kinds[k])
        n, nx, ny = request_shape(kind, int(sizes[k]))
        keps, mi = KIND_SETTINGS[kind]
        out.append(dict(kind=kind, dtype=0 if mix[2 * k + 1] < fp32_frac else 1, n=n, nx=nx,
                        ny=ny, seed=seed + ((k + 1) << 20), eps=keps if eps is None else eps,
                        max_iteration=mi,
                        arrival=float(arrivals[k])))
    return out


def problem_of(req: dict) -> Problem:
    """The host copy of a request's problem (the device generator builds the
    same arrays)."""
    kind, s = req["kind"], req["seed"]
    if kind == OBJ_GRID_QUAD:
        return grid_quadratic(req["nx"], req["ny"], seed=s)
    if kind == OBJ_GRID_COS:
        return grid_cos(req["nx"], req["ny"], seed=s)
    if kind == OBJ_ROSEN:
        return rosenbrock(req["n"], seed=s)
    return penta_qp(req["n"], seed=s)
