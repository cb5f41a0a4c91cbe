"""Serving loop over the C ABI (lbfgsb_serve_stream): a timed stream of
solve requests served with continuous batching under an HBM budget.

    reqs = problems.serving_stream(200, qps=5.0)
    out, stats = serve(reqs, budget_bytes=200e9)

`out` is a list of dicts (f, iter, info, gang, dtype, the arrival,
admission and completion times in seconds from the start of the stream, and
with polish=True whether the answer is the FP64 polish of an fp32 solve and
the fp32 phase's iterations); latency = t_done - t_arrival.  The polish
runs to the request's tolerance times polish_eps_scale (default 1e-2),
within the request's iteration cap.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib


def _torch():
    import torch
    if not torch.cuda.is_available():
        raise RuntimeError("lbfgsb_amd needs a ROCm GPU (torch.cuda.is_available() is False)")
    return torch


def serve(requests, budget_bytes: float = 200e9, fp32_share: float = 0.7, max_slots: int = 4096,
          threads: int = 256, blocks: int = 0, segment: int = 0, max_gang: int = 64,
          quantum: int = 20, use_graph: bool = True, polish: bool = False,
          polish_eps_scale: float = 1e-2):
    torch = _torch()
    L = _lib.lib()
    cnt = len(requests)
    arr = (_lib.Request * max(cnt, 1))()
    for i, r in enumerate(requests):
        q = arr[i]
        q.kind, q.dtype = int(r["kind"]), int(r["dtype"])
        q.n, q.nx, q.ny = int(r["n"]), int(r.get("nx", 0)), int(r.get("ny", 0))
        q.seed = int(r["seed"])
        q.eps = float(r["eps"])
        q.max_iteration = int(r["max_iteration"])
        q.arrival = float(r["arrival"])
    out = (_lib.Response * max(cnt, 1))()
    cfg = _lib.ServeConfig(hbm_budget_bytes=float(budget_bytes), fp32_share=float(fp32_share),
                           max_slots=int(max_slots), threads=int(threads), blocks=int(blocks),
                           segment=int(segment), max_gang=int(max_gang) if segment else 1,
                           quantum=int(quantum), use_graph=1 if use_graph else 0,
                           polish=1 if polish else 0,
                           polish_eps_scale=float(polish_eps_scale))
    st = _lib.ServeStats()
    stream = C.c_void_p(torch.cuda.current_stream().cuda_stream)
    rc = L.lbfgsb_serve_stream(C.byref(cfg), arr, cnt, out, C.byref(st), stream)
    if rc == -1:
        raise ValueError("lbfgsb_serve_stream: bad configuration")
    if rc != 0:
        raise RuntimeError(f"lbfgsb_serve_stream failed ({rc})")
    res = [dict(f=o.f, iter=o.iter, info=o.info, gang=o.gang, dtype=o.dtype,
                t_arrival=o.t_arrival, t_admit=o.t_admit, t_done=o.t_done,
                polished=o.polished, iter_fp32=o.iter_fp32)
           for o in out[:cnt]]
    stats = dict(seconds=st.seconds, launches=st.launches, arena_bytes=st.arena_bytes,
                 max_waiting=st.max_waiting, max_resident=st.max_resident)
    return res, stats


def latency_summary(res) -> dict:
    lat = np.array([r["t_done"] - r["t_arrival"] for r in res if r["info"] >= 0])
    if lat.size == 0:
        return {}
    return {"p50_ms": float(np.percentile(lat, 50) * 1e3),
            "p95_ms": float(np.percentile(lat, 95) * 1e3),
            "p99_ms": float(np.percentile(lat, 99) * 1e3),
            "mean_ms": float(lat.mean() * 1e3)}
