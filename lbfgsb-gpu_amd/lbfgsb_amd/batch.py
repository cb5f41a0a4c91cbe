"""Batched form over the C ABI (lbfgsb_batch_*): an HBM arena of problem
slots solved by the persistent per-problem kernel.

    b = Batch(np.float32, capacity=sum_n, max_problems=P)
    slots = [b.add(p, eps=1e-6, max_iter=1000) for p in problems]
    b.run()                     # every problem to completion
    f, iters, info, done = b.results()

`run(iter_budget=k)` advances each unfinished problem by at most k
iterations and returns how many remain; the serving loop admits new problems
between such calls.  `configure(segment, max_gang)` lets large problems be
solved by gangs of workgroups, one segment each.  `autotune()` picks the
workgroup size, the number of resident workgroups and the segment length by
timing whole solves of a sample of the workload.
"""
from __future__ import annotations

import ctypes as C
import time

import numpy as np

from . import _lib
from .problems import Problem


def _torch():
    import torch
    if not torch.cuda.is_available():
        raise RuntimeError("lbfgsb_amd needs a ROCm GPU (torch.cuda.is_available() is False)")
    return torch


class Batch:
    def __init__(self, dtype, capacity: int, max_problems: int, m: int = 8):
        self.torch = _torch()
        self.dtype = np.dtype(dtype)
        self.f64 = self.dtype == np.float64
        self.L = _lib.lib()
        self.stream = C.c_void_p(self.torch.cuda.current_stream().cuda_stream)
        self.h = self.L.lbfgsb_batch_create(1 if self.f64 else 0, int(capacity), int(max_problems),
                                            int(m), self.stream)
        if not self.h:
            raise RuntimeError("lbfgsb_batch_create failed (device memory?)")
        self.max_problems = int(max_problems)
        self.n = {}
        self.threads, self.blocks = 256, 0
        self.segment, self.max_gang = 0, 1

    def _dev(self, a, tdt):
        return self.torch.from_numpy(np.ascontiguousarray(a)).to("cuda", tdt)

    def add(self, p: Problem, eps: float = 1e-6, max_iter: int = 1000, eps_g=None, eps_f=None,
            eps_x=None) -> int:
        """Admit problem p (uploaded to the device first).  Returns the slot,
        -1 for bad input, -3 when the arena is full."""
        t = self.torch
        tdt = t.float64 if self.f64 else t.float32
        q = p.cast(self.dtype)
        arrs = {"x0": self._dev(q.x0, tdt), "l": self._dev(q.l, tdt), "u": self._dev(q.u, tdt),
                "nbd": self._dev(q.nbd, t.int32)}
        for k in ("b", "c", "a0", "a1", "a2"):
            a = getattr(q, k)
            if a is not None:
                arrs[k] = self._dev(a, tdt)
        ptr = lambda k: C.c_void_p(arrs[k].data_ptr()) if k in arrs else None
        slot = self.L.lbfgsb_batch_add(
            self.h, p.n, ptr("x0"), ptr("nbd"), ptr("l"), ptr("u"), p.kind, p.nx, p.ny, p.h,
            ptr("b"), ptr("c"), ptr("a0"), ptr("a1"), ptr("a2"),
            eps if eps_g is None else eps_g, eps if eps_f is None else eps_f,
            eps if eps_x is None else eps_x, int(max_iter))
        if slot >= 0:
            self.n[slot] = p.n
        return slot

    def add_device(self, n, x0, nbd, l, u, kind, nx=0, ny=0, h=0.0, b=None, c=None, a0=None,
                   a1=None, a2=None, eps=1e-6, max_iter=1000) -> int:
        """Admit a problem whose arrays are already device tensors."""
        ptr = lambda a: C.c_void_p(a.data_ptr()) if a is not None else None
        slot = self.L.lbfgsb_batch_add(self.h, int(n), ptr(x0), ptr(nbd), ptr(l), ptr(u),
                                       int(kind), int(nx), int(ny), float(h), ptr(b), ptr(c),
                                       ptr(a0), ptr(a1), ptr(a2), eps, eps, eps, int(max_iter))
        if slot >= 0:
            self.n[slot] = int(n)
        return slot

    def add_generated(self, kind: int, n: int, nx: int = 0, ny: int = 0, seed: int = 0,
                      eps: float = 1e-6, max_iter: int = 1000) -> int:
        """Admit a problem built on the device by the generator (the arrays
        problems.py builds for the same kind, shape and seed)."""
        slot = self.L.lbfgsb_batch_add_generated(self.h, int(kind), int(n), int(nx), int(ny),
                                                 int(seed), eps, eps, eps, int(max_iter))
        if slot >= 0:
            self.n[slot] = int(n)
        return slot

    def use_graph(self, on: bool = True):
        """Replay every launch from a captured HIP graph."""
        self.L.lbfgsb_batch_use_graph(self.h, 1 if on else 0)

    def run(self, iter_budget: int = 0, threads: int | None = None, blocks: int | None = None):
        th = self.threads if threads is None else threads
        bl = self.blocks if blocks is None else blocks
        left = self.L.lbfgsb_batch_run(self.h, int(th), int(bl), int(iter_budget))
        if left == -4:
            raise ValueError("a gang is wider than the launch (fewer workgroups than max_gang)")
        if left < 0:
            raise RuntimeError(f"lbfgsb_batch_run failed ({left})")
        return left

    def results(self):
        P = self.max_problems
        f = np.zeros(P, np.float64)
        it = np.zeros(P, np.int32)
        info = np.zeros(P, np.int32)
        done = np.zeros(P, np.int32)
        self.L.lbfgsb_batch_results(self.h, f.ctypes.data_as(C.c_void_p),
                                    it.ctypes.data_as(C.c_void_p),
                                    info.ctypes.data_as(C.c_void_p),
                                    done.ctypes.data_as(C.c_void_p))
        return f, it, info, done

    def clocks(self):
        """(start, end) device times in seconds of every slot's solve (end 0
        while unfinished)."""
        P = self.max_problems
        st = np.zeros(P, np.int64)
        en = np.zeros(P, np.int64)
        rate = C.c_double(0)
        self.L.lbfgsb_batch_clocks(self.h, st.ctypes.data_as(C.c_void_p),
                                   en.ctypes.data_as(C.c_void_p), C.byref(rate))
        return st / rate.value, en / rate.value

    def x(self, slot: int):
        t = self.torch
        out = t.empty(self.n[slot], dtype=t.float64 if self.f64 else t.float32, device="cuda")
        if self.L.lbfgsb_batch_get_x(self.h, slot, C.c_void_p(out.data_ptr())) != 0:
            raise RuntimeError("lbfgsb_batch_get_x failed")
        return out

    def remove(self, slot: int):
        if self.L.lbfgsb_batch_remove(self.h, slot) == 0:
            self.n.pop(slot, None)

    def free_elems(self) -> int:
        return int(self.L.lbfgsb_batch_free_elems(self.h))

    def configure(self, segment: int = 0, max_gang: int = 1):
        """Gangs: a problem of n elements is solved by min(max_gang,
        ceil(n / segment)) workgroups (segment >= 4096; 0 = one workgroup
        per problem)."""
        if self.L.lbfgsb_batch_configure(self.h, int(segment), int(max_gang)) != 0:
            raise ValueError(f"bad gang configuration segment={segment} max_gang={max_gang}")
        self.segment, self.max_gang = int(segment), int(max_gang)

    def gang(self, slot: int) -> int:
        return int(self.L.lbfgsb_batch_gang(self.h, slot))

    def close(self):
        if self.h:
            self.L.lbfgsb_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# (threads per workgroup, resident workgroups as a multiple of the CU count,
# segment length of the gangs (0 = one workgroup per problem))
TUNE_CANDIDATES = ((256, 2, 0), (512, 1, 0), (512, 1, 131072), (512, 1, 65536),
                   (512, 1, 32768), (512, 1, 16384), (256, 2, 65536), (256, 2, 32768))
MAX_GANG = 64


def add_any(b: Batch, item, eps=1e-6, max_iter=1000) -> int:
    """Admit a Problem (uploaded) or a request dict (generated on the
    device, with its own tolerance and iteration cap)."""
    if isinstance(item, Problem):
        return b.add(item, eps=eps, max_iter=max_iter)
    return b.add_generated(item["kind"], item["n"], item.get("nx", 0), item.get("ny", 0),
                           item["seed"], eps=item.get("eps", eps),
                           max_iter=item.get("max_iteration", max_iter))


def autotune(sample, dtype, eps=1e-6, max_iter=1000, candidates=TUNE_CANDIDATES,
             max_gang=MAX_GANG, graph=True):
    """Solve `sample` (Problems or request dicts drawn from the workload) once
    per candidate shape and return the fastest as {"threads", "blocks",
    "segment", "max_gang", ...}.  A shape that cannot hold the widest gang
    on one launch is skipped."""
    torch = _torch()
    cus = torch.cuda.get_device_properties(0).multi_processor_count
    cap = sum(p.n + 64 if isinstance(p, Problem) else p["n"] + 64 for p in sample)
    timings = []
    for th, per_cu, seg in candidates:
        b = Batch(dtype, cap, len(sample))
        b.configure(seg, max_gang if seg else 1)
        b.use_graph(graph)
        for p in sample:
            if add_any(b, p, eps, max_iter) < 0:
                raise RuntimeError("autotune: problem rejected")
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        try:
            b.run(threads=th, blocks=cus * per_cu)
        except ValueError:
            b.close()
            continue
        dt = time.perf_counter() - t0
        b.close()
        timings.append((dt, th, cus * per_cu, seg))
    dt, th, bl, seg = min(timings)
    return {"threads": th, "blocks": bl, "segment": seg, "max_gang": max_gang if seg else 1,
            "seconds": dt,
            "tried": [{"threads": t, "blocks": b_, "segment": g, "seconds": d}
                      for d, t, b_, g in timings]}
