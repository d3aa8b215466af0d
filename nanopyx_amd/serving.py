"""Served eSRRF request stream on one GPU: a request loop over the library's serving session (npx_session_*).

Requests arrive as a seeded Poisson process. Each is an eSRRF + temporal-correlation job (one timepoint of the
workflow, src/nanopyx/methods/esrrf/eSRRF_workflow.py:98-119) on its own stack, with its own frame count, frame size,
magnification, radius, sensitivity and temporal mode. The loop submits a request when its arrival time has passed;
the session admits it when its stack and running moments fit the device pool, and each step packs a frame slice of
every admitted job into shared launches, so requests that arrive while others are on the GPU join the work in
flight. Latency is arrival -> maps complete (observed by polling); throughput is output pixel-frames per second of
wall clock over the whole stream. ``serve_plain`` runs the same arrival trace as one plain npx_esrrf_temporal call
per request, first come first served, for comparison.
"""
from __future__ import annotations

import time
from dataclasses import dataclass, field

import numpy as np

from . import _hip, synthetic


@dataclass
class Request:
    stack: object                  # (N, H, W) torch tensor, float32 or uint16 (pinned host memory or device)
    magnification: int
    radius: float
    sensitivity: float
    mode: str                      # "AVG", "VAR" or "TAC2"
    iw: bool = True
    outputs: dict = field(default_factory=dict)

    @property
    def out_px_frames(self) -> int:
        N, H, W = self.stack.shape
        return int(N) * int(H) * self.magnification * int(W) * self.magnification


def make_requests(n, seed=5, host=True, max_frames=1000, device="cuda"):
    """n requests of the served-stream workload: frames log-uniform on 50-1000 (capped at max_frames), H and W from
    {64, 128, 256, 384, 512, 1024}, magnification from {2, 3, 4, 5, 8}, radius uniform on [1.0, 2.5], sensitivity
    from {1, 2, 3}, one of AVG / VAR / TAC2, float32 or uint16 camera counts, emitter density 0.02 per pixel.
    Stacks are rendered on the GPU; host=True moves them to pinned host memory (the session copies them in)."""
    import torch
    rng = np.random.Generator(np.random.PCG64(seed))
    reqs = []
    for i in range(n):
        sc = synthetic.request_scene(rng, seed * 100003 + i)
        sc.frames = min(sc.frames, max_frames)
        e = sc.esrrf
        x = synthetic.generate_device(sc, device=device)
        if host:
            h = torch.empty(x.shape, dtype=x.dtype, pin_memory=True)
            h.copy_(x)
            x = h
        reqs.append(Request(x, int(e["magnification"]), float(e["radius"]), float(e["sensitivity"]),
                            e["temporal"][0], True))
    if host:
        torch.cuda.synchronize()
    return reqs


def arrivals(n, rate_hz, seed=5):
    rng = np.random.Generator(np.random.PCG64(seed + 7))
    return np.cumsum(rng.exponential(1.0 / rate_hz, size=n))


def _alloc_outputs(torch, r: Request):
    N, H, W = r.stack.shape
    M = r.magnification
    r.outputs = {r.mode: torch.empty((H * M, W * M), dtype=torch.float32, device="cuda")}


def _summary(reqs, t_arr, t_done, wall):
    lat = np.asarray(t_done) - np.asarray(t_arr)
    px = sum(r.out_px_frames for r in reqs)
    return {"requests": len(reqs), "p50_ms": round(float(np.percentile(lat, 50)) * 1e3, 2),
            "p99_ms": round(float(np.percentile(lat, 99)) * 1e3, 2), "mean_ms": round(float(lat.mean()) * 1e3, 2),
            "max_ms": round(float(lat.max()) * 1e3, 2), "wall_s": round(wall, 3),
            "throughput_mpx_frames_s": round(px / 1e6 / wall, 1), "requests_per_s": round(len(reqs) / wall, 2)}


def by_size(reqs, lat, edges=None):
    """p50 / p99 latency (ms) of the requests split into three classes by output pixel-frames: below the first
    edge, between, above the second (edges default to the terciles of this request set)."""
    px = np.asarray([r.out_px_frames for r in reqs], np.float64)
    lat = np.asarray(lat, np.float64)
    if edges is None:
        edges = np.percentile(px, [100 / 3, 200 / 3])
    out = {}
    for name, sel in (("small", px <= edges[0]), ("medium", (px > edges[0]) & (px <= edges[1])),
                      ("large", px > edges[1])):
        if sel.any():
            out[name] = {"requests": int(sel.sum()), "max_mpx_frames": round(float(px[sel].max()) / 1e6, 1),
                         "p50_ms": round(float(np.percentile(lat[sel], 50)) * 1e3, 2),
                         "p99_ms": round(float(np.percentile(lat[sel], 99)) * 1e3, 2)}
    return out


def serve(reqs, rate_hz, seed=5, pool_bytes=64 << 30, max_active=32, px_frames_per_step=1 << 29, t_arr=None):
    """Runs the request stream through one session; returns (summary, per-request latencies in s)."""
    import torch
    t_arr = arrivals(len(reqs), rate_hz, seed) if t_arr is None else t_arr
    for r in reqs:
        _alloc_outputs(torch, r)
    torch.cuda.synchronize()
    sess = _hip.Session(pool_bytes, max_active, px_frames_per_step)
    stream = torch.cuda.current_stream()
    ids = {}
    t_done = [None] * len(reqs)
    nxt = 0
    t0 = time.perf_counter()
    try:
        while any(t is None for t in t_done):
            now = time.perf_counter() - t0
            while nxt < len(reqs) and t_arr[nxt] <= now:
                r = reqs[nxt]
                ids[nxt] = sess.submit(r.stack, r.magnification, r.radius, r.sensitivity, r.iw, (r.mode,), r.outputs)
                nxt += 1
            launched = sess.step(stream)
            finished = 0
            for i, jid in list(ids.items()):
                if sess.poll(jid) == _hip.Session.DONE:
                    t_done[i] = time.perf_counter() - t0
                    sess.forget(jid)
                    del ids[i]
                    finished += 1
            if not launched and not finished:
                time.sleep(2e-4)
        wall = time.perf_counter() - t0
        stats = sess.stats()
    finally:
        torch.cuda.synchronize()
        sess.close()
    out = _summary(reqs, t_arr, t_done, wall)
    out.update(launches=stats["launches"], slices=stats["slices"], steps=stats["steps"])
    return out, np.asarray(t_done) - np.asarray(t_arr)


def serve_plain(reqs, rate_hz, seed=5, t_arr=None):
    """The same arrival trace served one request at a time with one plain npx_esrrf_temporal call each (stack copied
    to the device first when it is on the host), first come first served."""
    import torch
    t_arr = arrivals(len(reqs), rate_hz, seed) if t_arr is None else t_arr
    t_done = [None] * len(reqs)
    t0 = time.perf_counter()
    for i, r in enumerate(reqs):
        now = time.perf_counter() - t0
        if t_arr[i] > now:
            time.sleep(t_arr[i] - now)
        x = r.stack if r.stack.is_cuda else r.stack.to("cuda", non_blocking=True)
        _hip.esrrf_temporal(x, r.magnification, r.radius, r.sensitivity, r.iw, (r.mode,))
        torch.cuda.synchronize()
        t_done[i] = time.perf_counter() - t0
    wall = time.perf_counter() - t0
    return _summary(reqs, t_arr, t_done, wall), np.asarray(t_done) - np.asarray(t_arr)
